// Internal declarations shared by the salfort translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/salfort.h"

namespace sal {

// ---------------------------------------------------------------------------------- errors
void set_error(const std::string& msg);

struct Error {
    int code;
    std::string msg;
};

#define SAL_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            throw ::sal::Error{SAL_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_) +   \
                                                " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"}; \
    } while (0)

#define SAL_REQUIRE(cond, code, msg)                                                               \
    do {                                                                                           \
        if (!(cond)) throw ::sal::Error{(code), std::string(msg)};                                 \
    } while (0)

// Runs f() and turns Error / std::bad_alloc / anything else into a status code.
template <class F>
int guard(F&& f) {
    try {
        f();
        return SAL_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return SAL_ERR_NOMEM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return SAL_ERR_STATE;
    }
}

// ---------------------------------------------------------------------------------- device memory
// Owning device buffer.  Allocation failures raise SAL_ERR_NOMEM.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            throw Error{SAL_ERR_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed: " +
                                           hipGetErrorString(e)};
        }
        n = count;
    }
    void ensure(size_t count) { if (count > n) alloc(count); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
    T* get() const { return p; }
};

template <class T>
void h2d(T* d, const T* h, size_t count, hipStream_t s) {
    if (count) SAL_HIP(hipMemcpyAsync(d, h, count * sizeof(T), hipMemcpyHostToDevice, s));
}
template <class T>
void d2h(T* h, const T* d, size_t count, hipStream_t s) {
    if (count) SAL_HIP(hipMemcpyAsync(h, d, count * sizeof(T), hipMemcpyDeviceToHost, s));
}

// Device-wide exclusive prefix sum of int32 (n >= 0), returns the total (synchronises `s`).
int64_t exclusive_scan_i32(const int* d_in, int* d_out, int64_t n, hipStream_t s);
// Stable radix sort of (key, value) int32 pairs on [0, 2^bits) keys.
void sort_pairs_i32(const int* kin, int* kout, const int* vin, int* vout, int64_t n, int bits, hipStream_t s);

// ---------------------------------------------------------------------------------- kernel timing
// When enabled (sal_prof_enable), every evaluation kernel launch is bracketed by HIP events on its
// own stream; sal_prof_get sums the durations per kernel name.
bool prof_on();
// Set while a launch sequence is being captured into a HIP graph on this thread (no timing events).
void set_capturing(bool on);
struct ProfScope {
    const char* name;
    hipStream_t s;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(const char* n, hipStream_t st);
    ~ProfScope();
};

// ---------------------------------------------------------------------------------- evaluation plans
struct Plan;       // per (degree, mode) evaluation plan, eval64.hip / evalmix.hip
void plan_free(Plan* p);

}  // namespace sal

// ---------------------------------------------------------------------------------- geometry handle
struct sal_geom_s {
    int64_t n = 0;
    int thresh = 0;
    hipStream_t stream = nullptr;  // the handle's own stream for set-up work

    // points in the caller's order
    sal::DBuf<double> x, y, z;
    sal::DBuf<double> xi, eta;  // face-local angles (device libm, then host-resolved near ties)
    sal::DBuf<int> pface;       // face 1..6
    sal::DBuf<int> anc;         // anc[L * n + p]: panel at level L containing p, -1 past its leaf
    int nlev = 0;               // levels in anc (max_level + 1)

    // panels: host mirror (authoritative for the panel-level bookkeeping) + device copies
    int P = 0;
    std::vector<int> level, is_leaf, parent, child1, face, count, level_lo;  // level_lo: BFS id range per level
    std::vector<double> bounds, radius, centre;  // 6 / 1 / 3 per panel
    std::vector<int64_t> start;                  // DFS start of each panel's point range
    sal::DBuf<int> d_level, d_is_leaf, d_parent, d_child1, d_face, d_count;
    sal::DBuf<double> d_bounds, d_radius, d_centre;
    sal::DBuf<int64_t> d_start;
    sal::DBuf<int> perm, iperm;              // DFS position -> caller index, and inverse
    sal::DBuf<double> px, py, pz;            // points in DFS order
    int64_t point_ties = 0, pair_ties = 0;

    // DTT list (reference order); list_thresh types its entries (DUAL_TREE_TRAVERSAL's
    // param_cluster_thresh, which need not be the tree's)
    int list_thresh = 0;
    int64_t nint = 0;
    int64_t type_count[4] = {0, 0, 0, 0};
    sal::DBuf<int> it, is, ty;

    // STT lists
    int64_t npc = 0, npp = 0;
    sal::DBuf<int> pc_t, pc_s, pp_t, pp_s;
    sal::DBuf<int> stt_counts;  // per target: pc count [n_own], then pp count [n_own]

    // one-interaction entry points (interact.hip): host copy of perm, device scratch
    std::vector<int> h_perm;
    sal::DBuf<double> one_buf;

    // cached evaluation plans (invalidated when the list changes); mu guards plan creation
    std::vector<sal::Plan*> plans;
    std::mutex mu;

    ~sal_geom_s();
};

namespace sal {
void invalidate_plans(sal_geom g);
void upload_panels(sal_geom g);  // device copies of the host panel arrays (d_level ... d_start)
void stt_release(sal_geom g);
// The STT targets and per-target offsets (kept with the handle in mom6.hip), copied out and restored
// by the geometry cache
struct SttState {
    int64_t nt = 0;
    std::vector<double> tx, ty, tz;
    std::vector<int> pc_off, pp_off;
};
void stt_get(sal_geom g, SttState& st);
void stt_set(sal_geom g, const SttState& st);
int64_t geom_device_bytes(sal_geom g);
}  // namespace sal
