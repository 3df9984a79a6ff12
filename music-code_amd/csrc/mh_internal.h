// Internal declarations of libmusic_hip (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/music_hip.h"

namespace mh {

// thread-local last error message, returned by mh_last_error()
void set_error(const std::string &msg);

#define MH_HIP(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            mh::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
            return MH_EHIP;                                                   \
        }                                                                     \
    } while (0)

#define MH_CHECK(cond, code, msg)  \
    do {                                 \
        if (!(cond)) {                   \
            mh::set_error(msg);          \
            return (code);               \
        }                                \
    } while (0)

// device buffer owned by the library (hipMalloc / hipFree)
struct DBuf {
    void *p = nullptr;
    size_t n = 0;
    ~DBuf() { release(); }
    int alloc(size_t bytes);
    void release();
    template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

constexpr int kMaxTypes = 64;

struct LJTable {
    int ntypes = 0;
    double A[kMaxTypes], B[kMaxTypes], hicut2[kMaxTypes], locut2[kMaxTypes];
    uint8_t on[kMaxTypes];
};

struct BuckTable {
    int ntypes = 0;
    double A[kMaxTypes], rho[kMaxTypes], C[kMaxTypes], D[kMaxTypes];
    double hicut2[kMaxTypes], locut2[kMaxTypes];
};

struct Ewald {
    bool set = false;
    double kappa_param = 0, kappa = 0, locut = 0;
    int kmax = 0, nk = 0;
    int sign_mode = 0;    // MH_EWALD_PARITY or MH_EWALD_PHYSICS
    DBuf k, s, si;        // accepted vectors, reference order
    DBuf ccoef;           // complex coefficient grid c(mx,my,mz) = s + i si, 0 if not kept
    std::vector<double> hk, hs, hsi;
};

}  // namespace mh

struct mh_ctx {
    int device = 0;
    // maps and GCMC batches keep their context alive: mh_ctx_destroy frees
    // it only when the last of them is destroyed
    int refs = 0;
    bool closing = false;
    hipStream_t stream = nullptr;
    // framework
    int natoms = 0;
    double fund_ell[3] = {0, 0, 0}, full_ell[3] = {0, 0, 0};
    // slanted cell (mh_framework_set_cell): unslant / slant matrices,
    // effective box of the maps (the fundamental cell edges when not set)
    bool slanted = false;
    double U[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, S[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double eff[3] = {0, 0, 0};
    double ang_deg[3] = {90, 90, 90};  // cell angles alpha, beta, gamma
    mh::DBuf xyz_s; // double [natoms][3] slanted atom coordinates S a
    mh::DBuf xyz;   // double [natoms][3]
    mh::DBuf type;  // int [natoms]
    mh::DBuf q;     // double [natoms]
    std::vector<double> hxyz, hq;
    std::vector<int> htype;
    int max_type = -1;
    mh::LJTable lj;
    bool lj_set = false;
    mh::BuckTable buck;
    bool buck_set = false;
    mh::Ewald ew;
    // scratch buffers reused across calls (grown on demand, freed with ctx)
    std::map<std::string, std::unique_ptr<mh::DBuf>> ws;
    // optional per-kernel timing with HIP events on the context stream
    bool timing = false;
    std::map<std::string, std::pair<double, long long>> ktimes;  // name -> (ms, launches)
};

namespace mh {
void ctx_retain(mh_ctx *ctx);
// uncompacted node-major copy of a map for the Monte Carlo lookups
// (built once per map, kept with it)
int map_dense(mh_ctx *ctx, mh_map *m, const float **out);
void ctx_release(mh_ctx *ctx);
// Scratch buffer 'name' of at least 'bytes' bytes owned by the context.
void *ws_get(mh_ctx *ctx, const std::string &name, size_t bytes, int *rc);
// once per device (mh_ctx_create): address of mh_map_build_static's scratch
int static_scratch_init(int device);

// Records events around a region on ctx->stream when timing is enabled
// (ctx may be NULL: no timing).
struct KTimer {
    mh_ctx *ctx;
    const char *name;
    hipEvent_t a = nullptr, b = nullptr;
    KTimer(mh_ctx *c, const char *n) : ctx(c), name(n) {
        if (ctx && ctx->timing) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, ctx->stream);
        }
    }
    void stop() {
        if (!ctx || !ctx->timing || !a) return;
        (void)hipEventRecord(b, ctx->stream);
        (void)hipEventSynchronize(b);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, a, b);
        auto &t = ctx->ktimes[name];
        t.first += ms;
        t.second += 1;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        a = b = nullptr;
    }
    ~KTimer() { stop(); }
};
}  // namespace mh

// bias map of biased insertion / deletion (mh_bmap.hip)
struct mh_bmap {
    mh_ctx *ctx = nullptr;
    int ncub[3] = {0, 0, 0};  // cubelets per axis
    long long ncubelets = 0, numtab = 0;
    double step[3] = {0, 0, 0}, eff[3] = {0, 0, 0};
    double cum_wt = 0.0, clamp_energy = 0.0;
    bool clamped = false;
    int gbits = 0;
    mh::DBuf cube, cindex, guide, cum, wt, wn;
};

struct mh_map {
    mh_ctx *ctx = nullptr;
    bool retained = false;
    int npnts[3] = {0, 0, 0};
    double step[3] = {0, 0, 0};
    double fund_ell[3] = {0, 0, 0};
    bool slanted = false;  // the framework's cell: maptouc in the slanted basis, no corner wrap
    double U[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, S[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    int64_t nfsize = 0;
    int64_t counts[4] = {0, 0, 0, 0};
    mh::DBuf ptr;    // int32 [nx][ny][nz] z fastest, 1-based
    mh::DBuf vals;   // float [nfsize][8] (AoS records, 32 B)
    mh::DBuf full32; // float [nodes][8] when kept
    mh::DBuf full64; // double [nodes][8] when kept
    // uncompacted node-major copy [nodes][8] for the GCMC lookups (one HBM
    // round trip per corner instead of ptr then value); V = NaN where the
    // node is untabulated.  Built on first use.
    mh::DBuf dense;
};
