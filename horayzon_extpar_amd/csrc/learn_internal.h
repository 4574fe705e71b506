// Objects of the learned elevation bracket shared by capi.cpp and learn.cpp.
#pragma once

#include "mlp.h"

struct hz_predictor {
    hz::Mlp net;
};
