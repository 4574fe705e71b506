# cython: language_level=3
"""The reference's Cython binding pattern, linked against libhorayzon.so.

The extern block declares horizon_svf_comp with the signature of the
reference's binding (horizon_svf.pyx:4-12, mo_lradtopo_horayzon.h:4-11),
taken from include/horayzon.h, so a maintainer's only edit to the
reference's .pyx is the header name.  horizon_svf_comp_relinked allocates
the outputs as the reference's wrapper does (horizon_svf.pyx:105-109:
horizon (num_hori, num_cell) and skyview (num_cell,), float64) and calls the
C symbol directly, with no checks of its own: the C library fills both with
NaN when it rejects the arguments.
"""

cimport numpy as cnp
import numpy as np

cnp.import_array()

cdef extern from "horayzon.h":
    void horizon_svf_comp(double* clon, double* clat, double* hsurf,
                          double* vlon, double* vlat,
                          cnp.npy_int32* cells_of_vertex,
                          double* horizon, double* skyview,
                          int num_cell, int num_vertex, int num_hori,
                          int grid_type, double dist_search,
                          double ray_org_elev, int refine_factor,
                          int svf_type)


def horizon_svf_comp_relinked(cnp.ndarray[cnp.float64_t, ndim=1, mode="c"] clon,
                              cnp.ndarray[cnp.float64_t, ndim=1, mode="c"] clat,
                              cnp.ndarray[cnp.float64_t, ndim=1, mode="c"] hsurf,
                              cnp.ndarray[cnp.float64_t, ndim=1, mode="c"] vlon,
                              cnp.ndarray[cnp.float64_t, ndim=1, mode="c"] vlat,
                              cnp.ndarray[cnp.int32_t, ndim=2, mode="c"] cells_of_vertex,
                              int num_hori, int grid_type, double dist_search,
                              double ray_org_elev, int refine_factor, int svf_type):
    """(horizon [deg], skyview [-]) from the C symbol horizon_svf_comp."""
    cdef int num_cell = clon.shape[0]
    cdef int num_vertex = vlon.shape[0]
    cdef cnp.ndarray[cnp.float64_t, ndim=2, mode="c"] horizon = \
        np.empty((num_hori, num_cell), dtype=np.float64)
    cdef cnp.ndarray[cnp.float64_t, ndim=1, mode="c"] skyview = \
        np.empty(num_cell, dtype=np.float64)
    horizon_svf_comp(&clon[0], &clat[0], &hsurf[0], &vlon[0], &vlat[0],
                     <cnp.npy_int32*>&cells_of_vertex[0, 0], &horizon[0, 0], &skyview[0],
                     num_cell, num_vertex, num_hori, grid_type, dist_search, ray_org_elev,
                     refine_factor, svf_type)
    return horizon, skyview
