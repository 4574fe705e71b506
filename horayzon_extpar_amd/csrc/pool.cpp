// Persistent host worker pool of the library's host-side loops (the FP64
// prologue passes and the pinned-memory staging copies).  Creating and
// joining a thread per pass cost ~0.5 ms per pass with 8-16 threads, more
// than some of the passes themselves.
//
// Tasks never wait on other tasks: a caller submits plain work items to a
// group and then waits for the group, running queued items (its own or
// another caller's) while it waits.  Calls from several host threads then
// share the workers and always finish.  A forked child process gets a pool
// of its own on first use (the parent's worker threads do not exist there).

#include <unistd.h>

#include <algorithm>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>

#include "hz_internal.h"

namespace hz {

namespace {

struct Pool {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::pair<TaskGroup*, std::function<void()>>> q;
    int workers = 0;
    pid_t pid = 0;

    // pops and runs one queued item; false when the queue is empty
    bool run_one() {
        std::pair<TaskGroup*, std::function<void()>> t;
        {
            std::lock_guard<std::mutex> g(mu);
            if (q.empty()) return false;
            t = std::move(q.front());
            q.pop_front();
        }
        t.second();
        t.first->left.fetch_sub(1, std::memory_order_acq_rel);
        return true;
    }

    void loop() {
        while (true) {
            std::pair<TaskGroup*, std::function<void()>> t;
            {
                std::unique_lock<std::mutex> g(mu);
                cv.wait(g, [this] { return !q.empty(); });
                t = std::move(q.front());
                q.pop_front();
            }
            t.second();
            t.first->left.fetch_sub(1, std::memory_order_acq_rel);
        }
    }
};

std::mutex g_pool_mu;
Pool* g_pool = nullptr;

// The pool of this process, created on first use with one worker per
// available CPU but one (the caller works too), at most 15.  Pools are never
// destroyed: idle workers block on the condition variable until exit.
Pool& pool() {
    std::lock_guard<std::mutex> g(g_pool_mu);
    const pid_t me = getpid();
    if (!g_pool || g_pool->pid != me) {
        Pool* p = new Pool();
        p->pid = me;
        unsigned hw = std::thread::hardware_concurrency();
        const int want = (int)std::min(hw > 1 ? hw - 1 : 0u, 15u);
        for (int i = 0; i < want; i++) {
            try {
                std::thread(&Pool::loop, p).detach();
                p->workers++;
            } catch (...) {
                break;  // fewer workers; the callers run the rest themselves
            }
        }
        g_pool = p;  // a pool of a parent process is left behind in a fork child
    }
    return *g_pool;
}

}  // namespace

int pool_threads() { return pool().workers + 1; }

void pool_submit(TaskGroup& grp, std::function<void()> f) {
    Pool& p = pool();
    grp.left.fetch_add(1, std::memory_order_acq_rel);
    {
        std::lock_guard<std::mutex> g(p.mu);
        p.q.emplace_back(&grp, std::move(f));
    }
    p.cv.notify_one();
}

void pool_wait(TaskGroup& grp) {
    Pool& p = pool();
    while (grp.left.load(std::memory_order_acquire) > 0)
        if (!p.run_one()) std::this_thread::yield();
}

void pool_for(long long n, const std::function<void(long long, long long)>& f,
              long long min_per_task) {
    if (n <= 0) return;
    const int nt = (int)std::max(
        1LL, std::min<long long>((long long)pool_threads(), n / std::max(1LL, min_per_task)));
    if (nt == 1) {
        f(0, n);
        return;
    }
    TaskGroup grp;
    const long long chunk = (n + nt - 1) / nt;
    for (int t = 1; t < nt; t++) {
        const long long b = t * chunk, e = std::min(n, b + chunk);
        if (b < e) pool_submit(grp, [&f, b, e]() { f(b, e); });
    }
    f(0, std::min(n, chunk));
    pool_wait(grp);
}

}  // namespace hz
