// libmusic_hip: context, framework/forcefield upload, Ewald structure factor.
#include <cmath>
#include <cstring>
#include <cstdio>

#include "mh_internal.h"

namespace mh {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

int DBuf::alloc(size_t bytes) {
    release();
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        p = nullptr;
        set_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
        return MH_ENOMEM;
    }
    n = bytes;
    return MH_OK;
}

void DBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
}

void ctx_retain(mh_ctx *ctx) { ctx->refs++; }

void ctx_release(mh_ctx *ctx) {
    if (--ctx->refs == 0 && ctx->closing) mh_ctx_destroy(ctx);
}

void *ws_get(mh_ctx *ctx, const std::string &name, size_t bytes, int *rc) {
    auto &slot = ctx->ws[name];
    if (!slot) slot.reset(new DBuf());
    if (slot->n < bytes) {
        (void)hipStreamSynchronize(ctx->stream);  // no kernel may still use the old buffer
        size_t grow = bytes + bytes / 8;
        int r = slot->alloc(grow);
        if (r) {
            *rc = r;
            return nullptr;
        }
    }
    *rc = MH_OK;
    return slot->p;
}

// Reference constants (defaults.F90:33-35): truncated pi.
constexpr double kPi = 3.1415927;
constexpr double kTwoPi = 2.0 * kPi;

// ewald_sfactor / ewald_sfSum (ewald.F90:558-681): one lane per candidate
// k = (x, y, z) in the reference loop order; the structure-factor sums run
// over the atoms in their given order, exactly like the reference loop.
__global__ void k_sfactor(int kmax, int natoms, const double *__restrict__ xyz,
                          const double *__restrict__ q, double ialpha, double ibeta,
                          double igamma, double ikapPiSq, double iVolPi, double locut,
                          uint8_t *__restrict__ keep, double *__restrict__ outs,
                          double *__restrict__ outsi, double *__restrict__ outk) {
    const int side = 2 * kmax + 1;
    const int ncand = side * side * side;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncand) return;
    int x = c / (side * side) - kmax;
    int y = (c / side) % side - kmax;
    int z = c % side - kmax;
    keep[c] = 0;
    if (x == 0 && y == 0 && z == 0) return;
    double kalpha = (double)x * ialpha, kbeta = (double)y * ibeta, kgamma = (double)z * igamma;
    double ka2pi = kalpha * kTwoPi, kb2pi = kbeta * kTwoPi, kg2pi = kgamma * kTwoPi;
    double ksqx = kalpha * kalpha, ksqy = kbeta * kbeta, ksqz = kgamma * kgamma;
    double ksq = __dadd_rn(__dadd_rn(ksqx, ksqy), ksqz);
    double factor = __ddiv_rn(__dmul_rn(iVolPi, exp(-ikapPiSq * ksq)), ksq);
    if (!(factor > locut)) return;
    double sum = 0.0, sumi = 0.0;
    for (int i = 0; i < natoms; i++) {
        double rdotk = __dadd_rn(__dadd_rn(__dmul_rn(ka2pi, xyz[3 * i]), __dmul_rn(kb2pi, xyz[3 * i + 1])),
                                 __dmul_rn(kg2pi, xyz[3 * i + 2]));
        double sn, cs;
        sincos(rdotk, &sn, &cs);
        sum = __dadd_rn(sum, __dmul_rn(cs, q[i]));
        sumi = __dadd_rn(sumi, __dmul_rn(sn, q[i]));
    }
    double smag = __dmul_rn(__dmul_rn(factor, factor), __dadd_rn(__dmul_rn(sum, sum), __dmul_rn(sumi, sumi)));
    if (smag <= locut * 1.0e-3) return;
    keep[c] = 1;
    outs[c] = sum * factor;
    outsi[c] = sumi * factor;
    outk[3 * c] = ka2pi;
    outk[3 * c + 1] = kb2pi;
    outk[3 * c + 2] = kg2pi;
}

}  // namespace mh

using namespace mh;

extern "C" {

const char *mh_last_error(void) { return g_err.c_str(); }
const char *mh_version(void) { return "music_hip 0.1 gfx950"; }

int mh_ctx_create(int device, mh_ctx **out) {
    MH_CHECK(out, MH_EINVAL, "mh_ctx_create: out is NULL");
    int ndev = 0;
    MH_HIP(hipGetDeviceCount(&ndev));
    MH_CHECK(device >= 0 && device < ndev, MH_EINVAL, "mh_ctx_create: no such device");
    MH_HIP(hipSetDevice(device));
    // device address of the map builder's static scratch on this device; a
    // failure is recorded there and only mh_map_build_static refuses to run
    (void)static_scratch_init(device);
    mh_ctx *c = new mh_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
        return MH_EHIP;
    }
    *out = c;
    return MH_OK;
}

int mh_ctx_destroy(mh_ctx *ctx) {
    if (!ctx) return MH_OK;
    ctx->closing = true;
    if (ctx->refs == 0) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        ctx->ws.clear();
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    }
    return MH_OK;
}

static int check_cell(const double e[3]) {
    for (int i = 0; i < 3; i++)
        if (!(e[i] > 0.0) || !std::isfinite(e[i])) return 0;
    return 1;
}

int mh_framework_upload(mh_ctx *ctx, int natoms, const double *xyz, const int *type,
                        const double *q, const double fund_ell[3], const double full_ell[3]) {
    MH_CHECK(ctx, MH_EINVAL, "mh_framework_upload: ctx is NULL");
    MH_CHECK(natoms >= 0 && (natoms == 0 || (xyz && type)) && fund_ell && full_ell, MH_EINVAL,
             "mh_framework_upload: bad arguments");
    MH_CHECK(check_cell(fund_ell) && check_cell(full_ell), MH_EINVAL,
             "mh_framework_upload: cell edges must be positive");
    MH_HIP(hipSetDevice(ctx->device));
    ctx->natoms = natoms;
    ctx->hxyz.assign(xyz, xyz + 3 * (size_t)natoms);
    ctx->htype.assign(type, type + natoms);
    ctx->hq.assign(natoms, 0.0);
    if (q) ctx->hq.assign(q, q + natoms);
    ctx->max_type = -1;
    for (int i = 0; i < natoms; i++) {
        MH_CHECK(type[i] >= 0 && type[i] < kMaxTypes, MH_EINVAL,
                 "mh_framework_upload: atom type out of range [0,64)");
        if (type[i] > ctx->max_type) ctx->max_type = type[i];
    }
    for (int i = 0; i < 3; i++) {
        ctx->fund_ell[i] = fund_ell[i];
        ctx->full_ell[i] = full_ell[i];
        ctx->eff[i] = fund_ell[i];
        ctx->ang_deg[i] = 90.0;
    }
    ctx->slanted = false;
    ctx->ew.set = false;
    if (natoms == 0) return MH_OK;  // a cell without framework atoms (sorbate-only systems)
    int rc;
    if ((rc = ctx->xyz.alloc(sizeof(double) * 3 * natoms))) return rc;
    if ((rc = ctx->type.alloc(sizeof(int) * natoms))) return rc;
    if ((rc = ctx->q.alloc(sizeof(double) * natoms))) return rc;
    MH_HIP(hipMemcpy(ctx->xyz.p, ctx->hxyz.data(), sizeof(double) * 3 * natoms, hipMemcpyHostToDevice));
    MH_HIP(hipMemcpy(ctx->type.p, type, sizeof(int) * natoms, hipMemcpyHostToDevice));
    MH_HIP(hipMemcpy(ctx->q.p, ctx->hq.data(), sizeof(double) * natoms, hipMemcpyHostToDevice));
    return MH_OK;
}

// matrix_inverse (matrix.F90:195-229): LU with implicit scaling and partial
// pivoting, one back substitution per column (matrixops_ludcmpReal /
// lubksbReal, matrixops.F90:31-161); row-major 3x3
static void lu_inverse3(const double m[9], double inv[9]) {
    double a[3][3], vv[3];
    int indx[3];
    for (int i = 0; i < 9; i++) a[i / 3][i % 3] = m[i];
    for (int i = 0; i < 3; i++) {
        double aamax = 0.0;
        for (int j = 0; j < 3; j++)
            if (std::fabs(a[i][j]) > aamax) aamax = std::fabs(a[i][j]);
        vv[i] = 1.0 / aamax;
    }
    for (int j = 0; j < 3; j++) {
        for (int i = 1; i < j; i++) {
            double sum = a[i][j];
            for (int k = 0; k < i; k++) sum = sum - a[i][k] * a[k][j];
            a[i][j] = sum;
        }
        double aamax = 0.0;
        int imax = j;
        for (int i = j; i < 3; i++) {
            double sum = a[i][j];
            if (j > 0) {
                for (int k = 0; k < j; k++) sum = sum - a[i][k] * a[k][j];
                a[i][j] = sum;
            }
            double dum = vv[i] * std::fabs(sum);
            if (dum >= aamax) {
                imax = i;
                aamax = dum;
            }
        }
        if (j != imax) {
            for (int k = 0; k < 3; k++) std::swap(a[imax][k], a[j][k]);
            vv[imax] = vv[j];
        }
        indx[j] = imax;
        if (j != 2) {
            if (a[j][j] == 0.0) a[j][j] = (double)1.0e-20f;
            double dum = 1.0 / a[j][j];
            for (int i = j + 1; i < 3; i++) a[i][j] = a[i][j] * dum;
        }
    }
    if (a[2][2] == 0.0) a[2][2] = (double)1.0e-20f;
    for (int col = 0; col < 3; col++) {
        double b[3] = {0.0, 0.0, 0.0};
        b[col] = 1.0;
        int ii = -1;
        for (int i = 0; i < 3; i++) {
            const int ll = indx[i];
            double sum = b[ll];
            b[ll] = b[i];
            if (ii >= 0) {
                for (int j = ii; j < i; j++) sum = sum - a[i][j] * b[j];
            } else if (sum != 0.0) {
                ii = i;
            }
            b[i] = sum;
        }
        for (int i = 2; i >= 0; i--) {
            double sum = b[i];
            for (int j = i + 1; j < 3; j++) sum = sum - a[i][j] * b[j];
            b[i] = sum / a[i][i];
        }
        for (int i = 0; i < 3; i++) inv[3 * i + col] = b[i];
    }
}

int mh_framework_set_cell(mh_ctx *ctx, const double angles_deg[3], const double eff[3]) {
    MH_CHECK(ctx && angles_deg && eff, MH_EINVAL, "mh_framework_set_cell: bad arguments");
    MH_CHECK(check_cell(eff), MH_EINVAL, "mh_framework_set_cell: effective box edges must be positive");
    for (int i = 0; i < 3; i++)
        MH_CHECK(angles_deg[i] > 0.0 && angles_deg[i] < 180.0, MH_EINVAL,
                 "mh_framework_set_cell: angles must lie in (0, 180) degrees");
    // fundcell_init (fundcell.F90:60-112): orthorhombic within 0.001 degree
    const double err = std::fabs(angles_deg[0] - 90.0) + std::fabs(angles_deg[1] - 90.0) +
                       std::fabs(angles_deg[2] - 90.0);
    for (int i = 0; i < 3; i++) {
        ctx->eff[i] = eff[i];
        ctx->ang_deg[i] = angles_deg[i];
    }
    ctx->slanted = err > 0.001;
    if (!ctx->slanted) return MH_OK;
    // fundcell_initxform (fundcell.F90:378-412): the angles in default real,
    // single-precision cosines
    const double degtorad = 3.1415927 / 180.0;
    const float alpha = (float)(angles_deg[0] * degtorad), beta = (float)(angles_deg[1] * degtorad),
                gamma = (float)(angles_deg[2] * degtorad);
    const float cg = cosf(gamma), cb = cosf(beta), ca = cosf(alpha);
    double U[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    U[0] = 1.0;
    U[1] = (double)cg;
    U[2] = (double)cb;
    const float cg2 = cg * cg;
    U[4] = std::sqrt(1.0 - (double)cg2);
    const float num = ca - cg * cb;
    U[5] = (double)num / U[4];
    U[8] = std::sqrt(1.0 - U[2] * U[2] - U[5] * U[5]);
    MH_CHECK(std::isfinite(U[5]) && std::isfinite(U[8]) && U[8] > 0.0, MH_EINVAL,
             "mh_framework_set_cell: the angles do not form a cell");
    double S[9];
    lu_inverse3(U, S);
    for (int i = 0; i < 9; i++) {
        ctx->U[i] = U[i];
        ctx->S[i] = S[i];
    }
    // slanted atom coordinates S a (fundcell_slant, matrix_m_mult_v order)
    if (ctx->natoms > 0) {
        std::vector<double> xs(3 * (size_t)ctx->natoms);
        for (int a = 0; a < ctx->natoms; a++)
            for (int r = 0; r < 3; r++) {
                double s = 0.0;
                for (int c = 0; c < 3; c++) s = s + S[3 * r + c] * ctx->hxyz[3 * a + c];
                xs[3 * a + r] = s;
            }
        int rc;
        if ((rc = ctx->xyz_s.alloc(sizeof(double) * xs.size()))) return rc;
        MH_HIP(hipSetDevice(ctx->device));
        MH_HIP(hipMemcpy(ctx->xyz_s.p, xs.data(), sizeof(double) * xs.size(), hipMemcpyHostToDevice));
    }
    return MH_OK;
}

int mh_ff_set_lj(mh_ctx *ctx, int ntypes, const double *A, const double *B,
                 const double *hicut2, const double *locut2, const uint8_t *on) {
    MH_CHECK(ctx && A && B && hicut2 && locut2, MH_EINVAL, "mh_ff_set_lj: bad arguments");
    MH_CHECK(ntypes > 0 && ntypes <= kMaxTypes, MH_EINVAL, "mh_ff_set_lj: ntypes out of range");
    LJTable &t = ctx->lj;
    t.ntypes = ntypes;
    for (int i = 0; i < ntypes; i++) {
        t.A[i] = A[i];
        t.B[i] = B[i];
        t.hicut2[i] = hicut2[i];
        t.locut2[i] = locut2[i];
        t.on[i] = on ? on[i] : 1;
        if (!t.on[i]) t.A[i] = t.B[i] = t.hicut2[i] = t.locut2[i] = 0.0;
    }
    ctx->lj_set = true;
    return MH_OK;
}

int mh_ff_set_buck(mh_ctx *ctx, int ntypes, const double *A, const double *rho,
                   const double *C, const double *D, const double *hicut2,
                   const double *locut2) {
    MH_CHECK(ctx && A && rho && C && D && hicut2 && locut2, MH_EINVAL,
             "mh_ff_set_buck: bad arguments");
    MH_CHECK(ntypes > 0 && ntypes <= kMaxTypes, MH_EINVAL, "mh_ff_set_buck: ntypes out of range");
    BuckTable &t = ctx->buck;
    t.ntypes = ntypes;
    for (int i = 0; i < ntypes; i++) {
        MH_CHECK(rho[i] > 0.0 || A[i] == 0.0, MH_EINVAL, "mh_ff_set_buck: rho must be > 0");
        t.A[i] = A[i];
        t.rho[i] = rho[i] > 0.0 ? rho[i] : 1.0;
        t.C[i] = C[i];
        t.D[i] = D[i];
        t.hicut2[i] = hicut2[i];
        t.locut2[i] = locut2[i];
    }
    ctx->buck_set = true;
    return MH_OK;
}

int mh_ewald_set(mh_ctx *ctx, double kappa_param, int kmax, double locut, int sign_mode, int *nk_out) {
    MH_CHECK(ctx, MH_EINVAL, "mh_ewald_set: ctx is NULL");
    MH_CHECK(sign_mode == MH_EWALD_PARITY || sign_mode == MH_EWALD_PHYSICS, MH_EINVAL,
             "mh_ewald_set: sign_mode must be MH_EWALD_PARITY or MH_EWALD_PHYSICS");
    MH_CHECK(ctx->natoms > 0, MH_ESTATE, "mh_ewald_set: upload the framework first");
    MH_CHECK(kmax >= 1 && kmax <= 40, MH_EINVAL, "mh_ewald_set: kmax out of range [1,40]");
    MH_CHECK(kappa_param > 0.0, MH_EINVAL, "mh_ewald_set: kappa must be > 0");
    MH_HIP(hipSetDevice(ctx->device));
    Ewald &ew = ctx->ew;
    const double *ell = ctx->full_ell;
    double mn = ell[0] < ell[1] ? ell[0] : ell[1];
    mn = mn < ell[2] ? mn : ell[2];
    ew.kappa_param = kappa_param;
    ew.kappa = kappa_param / mn;
    ew.kmax = kmax;
    ew.locut = locut;
    ew.sign_mode = sign_mode;
    double ialpha = 1.0 / ell[0], ibeta = 1.0 / ell[1], igamma = 1.0 / ell[2];
    double ikap = 1.0 / ew.kappa;
    double pk = kPi * ikap;
    double ikapPiSq = pk * pk;
    double iVolPi = ialpha * ibeta * igamma / kPi;
    const int side = 2 * kmax + 1;
    const int ncand = side * side * side;
    DBuf keep, cs, csi, ck;
    int rc;
    if ((rc = keep.alloc(ncand))) return rc;
    if ((rc = cs.alloc(sizeof(double) * ncand))) return rc;
    if ((rc = csi.alloc(sizeof(double) * ncand))) return rc;
    if ((rc = ck.alloc(sizeof(double) * 3 * ncand))) return rc;
    int bs = 64;
    hipLaunchKernelGGL(k_sfactor, dim3((ncand + bs - 1) / bs), dim3(bs), 0, ctx->stream, kmax,
                       ctx->natoms, ctx->xyz.as<double>(), ctx->q.as<double>(), ialpha, ibeta,
                       igamma, ikapPiSq, iVolPi, locut, keep.as<uint8_t>(), cs.as<double>(),
                       csi.as<double>(), ck.as<double>());
    MH_HIP(hipGetLastError());
    std::vector<uint8_t> hkeep(ncand);
    std::vector<double> hs(ncand), hsi(ncand), hk(3 * (size_t)ncand);
    MH_HIP(hipMemcpyAsync(hkeep.data(), keep.p, ncand, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipMemcpyAsync(hs.data(), cs.p, sizeof(double) * ncand, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipMemcpyAsync(hsi.data(), csi.p, sizeof(double) * ncand, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipMemcpyAsync(hk.data(), ck.p, sizeof(double) * 3 * ncand, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    // ordered compaction of the accepted vectors (reference order x,y,z)
    ew.hk.clear();
    ew.hs.clear();
    ew.hsi.clear();
    std::vector<double> grid(2 * (size_t)ncand, 0.0);  // complex c = s + i si per candidate
    for (int c = 0; c < ncand; c++) {
        if (!hkeep[c]) continue;
        ew.hs.push_back(hs[c]);
        ew.hsi.push_back(hsi[c]);
        ew.hk.push_back(hk[3 * c]);
        ew.hk.push_back(hk[3 * c + 1]);
        ew.hk.push_back(hk[3 * c + 2]);
        grid[2 * c] = hs[c];
        grid[2 * c + 1] = hsi[c];
    }
    ew.nk = (int)ew.hs.size();
    MH_CHECK(ew.nk <= 50000, MH_ECAPACITY,
             "mh_ewald_set: more than nkmax = 50000 k vectors (ewald.F90:134)");
    if ((rc = ew.k.alloc(sizeof(double) * 3 * (ew.nk + 1)))) return rc;
    if ((rc = ew.s.alloc(sizeof(double) * (ew.nk + 1)))) return rc;
    if ((rc = ew.si.alloc(sizeof(double) * (ew.nk + 1)))) return rc;
    if ((rc = ew.ccoef.alloc(sizeof(double) * 2 * ncand))) return rc;
    if (ew.nk) {
        MH_HIP(hipMemcpy(ew.k.p, ew.hk.data(), sizeof(double) * 3 * ew.nk, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(ew.s.p, ew.hs.data(), sizeof(double) * ew.nk, hipMemcpyHostToDevice));
        MH_HIP(hipMemcpy(ew.si.p, ew.hsi.data(), sizeof(double) * ew.nk, hipMemcpyHostToDevice));
    }
    MH_HIP(hipMemcpy(ew.ccoef.p, grid.data(), sizeof(double) * 2 * ncand, hipMemcpyHostToDevice));
    ew.set = true;
    if (nk_out) *nk_out = ew.nk;
    return MH_OK;
}

int mh_ctx_set_timing(mh_ctx *ctx, int on) {
    MH_CHECK(ctx, MH_EINVAL, "mh_ctx_set_timing: ctx is NULL");
    ctx->timing = on != 0;
    ctx->ktimes.clear();
    return MH_OK;
}

int mh_ctx_get_timing(mh_ctx *ctx, char *buf, int len) {
    MH_CHECK(ctx && buf && len > 0, MH_EINVAL, "mh_ctx_get_timing: bad arguments");
    std::string js = "{";
    bool first = true;
    for (auto &kv : ctx->ktimes) {
        char tmp[256];
        snprintf(tmp, sizeof tmp, "%s\"%s\": [%.6f, %lld]", first ? "" : ", ", kv.first.c_str(),
                 kv.second.first, kv.second.second);
        js += tmp;
        first = false;
    }
    js += "}";
    MH_CHECK((int)js.size() < len, MH_ECAPACITY, "mh_ctx_get_timing: buffer too small");
    memcpy(buf, js.c_str(), js.size() + 1);
    return MH_OK;
}

int mh_ewald_get(mh_ctx *ctx, double *k, double *s, double *si, double *kappa) {
    MH_CHECK(ctx && ctx->ew.set, MH_ESTATE, "mh_ewald_get: call mh_ewald_set first");
    const Ewald &ew = ctx->ew;
    if (k) std::memcpy(k, ew.hk.data(), sizeof(double) * 3 * ew.nk);
    if (s) std::memcpy(s, ew.hs.data(), sizeof(double) * ew.nk);
    if (si) std::memcpy(si, ew.hsi.data(), sizeof(double) * ew.nk);
    if (kappa) *kappa = ew.kappa;
    return MH_OK;
}

}  // extern "C"
