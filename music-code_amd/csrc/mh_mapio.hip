// libmusic_hip: map object access, map file I/O and map interpolation.
#include <cstdio>
#include <cstring>
#include <vector>

#include "mh_device.h"
#include "mh_internal.h"

namespace mh {

__global__ void k_interp(int n, const double *__restrict__ xyz, const int32_t *__restrict__ ptr,
                         const float *__restrict__ vals, int nx, int ny, int nz, double sx,
                         double sy, double sz, double ex, double ey, double ez, double factor,
                         int slanted, SlantCell sc, double *__restrict__ e, uint8_t *__restrict__ ok) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double p;
    bool good = hermite_point(ptr, vals, nx, ny, nz, sx, sy, sz, ex, ey, ez, xyz[3 * i],
                              xyz[3 * i + 1], xyz[3 * i + 2], p, slanted ? &sc : nullptr);
    e[i] = good ? p * factor : 0.0;
    ok[i] = good ? 1 : 0;
}

// SoA [8][nf] <-> AoS [nf][8]
__global__ void k_aos_to_soa(long long nf, const float *__restrict__ aos, float *__restrict__ soa) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    for (int a = 0; a < 8; a++) soa[a * nf + i] = aos[8 * i + a];
}

__global__ void k_soa_to_aos(long long nf, const float *__restrict__ soa, float *__restrict__ aos) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    for (int a = 0; a < 8; a++) aos[8 * i + a] = soa[a * nf + i];
}

// Fortran unformatted sequential record I/O (4-byte markers)
static bool rec_write(FILE *f, const void *buf, size_t n) {
    if (n > 0x7fffffffULL) return false;
    uint32_t m = (uint32_t)n;
    return fwrite(&m, 4, 1, f) == 1 && (n == 0 || fwrite(buf, 1, n, f) == n) && fwrite(&m, 4, 1, f) == 1;
}

static bool rec_read(FILE *f, void *buf, size_t cap, size_t *got) {
    uint32_t m1, m2;
    if (fread(&m1, 4, 1, f) != 1 || m1 > cap) return false;
    if (m1 && fread(buf, 1, m1, f) != m1) return false;
    if (fread(&m2, 4, 1, f) != 1 || m1 != m2) return false;
    if (got) *got = m1;
    return true;
}

static void padcpy(char *dst, const char *src, size_t n) {
    size_t l = 0;
    while (l < n && src[l]) l++;
    memset(dst, ' ', n);
    memcpy(dst, src, l);
}

}  // namespace mh

using namespace mh;

extern "C" {

int mh_map_destroy(mh_map *map) {
    if (!map) return MH_OK;
    mh_ctx *ctx = map->ctx;
    bool ret = map->retained;
    if (ctx) (void)hipSetDevice(ctx->device);
    delete map;
    if (ctx && ret) ctx_release(ctx);
    return MH_OK;
}

int mh_map_info(const mh_map *map, int npnts[3], int64_t *nfsize, int64_t counts[4], double step[3]) {
    MH_CHECK(map, MH_EINVAL, "mh_map_info: map is NULL");
    for (int i = 0; i < 3; i++) {
        if (npnts) npnts[i] = map->npnts[i];
        if (step) step[i] = map->step[i];
    }
    if (nfsize) *nfsize = map->nfsize;
    if (counts)
        for (int i = 0; i < 4; i++) counts[i] = map->counts[i];
    return MH_OK;
}

int mh_map_get(const mh_map *map, int32_t *ptr, float *vals8) {
    MH_CHECK(map, MH_EINVAL, "mh_map_get: map is NULL");
    mh_ctx *ctx = map->ctx;
    MH_HIP(hipSetDevice(ctx->device));
    const long long nodes = (long long)map->npnts[0] * map->npnts[1] * map->npnts[2];
    if (ptr) MH_HIP(hipMemcpy(ptr, map->ptr.p, sizeof(int32_t) * nodes, hipMemcpyDeviceToHost));
    if (vals8 && map->nfsize > 0) {
        DBuf soa;
        int rc;
        if ((rc = soa.alloc(sizeof(float) * 8 * map->nfsize))) return rc;
        hipLaunchKernelGGL(k_aos_to_soa, dim3((unsigned)((map->nfsize + 255) / 256)), dim3(256), 0,
                           ctx->stream, (long long)map->nfsize, map->vals.as<float>(), soa.as<float>());
        MH_HIP(hipGetLastError());
        MH_HIP(hipStreamSynchronize(ctx->stream));
        MH_HIP(hipMemcpy(vals8, soa.p, sizeof(float) * 8 * map->nfsize, hipMemcpyDeviceToHost));
    }
    return MH_OK;
}

int mh_map_get_full(const mh_map *map, float *f32, double *f64) {
    MH_CHECK(map, MH_EINVAL, "mh_map_get_full: map is NULL");
    MH_CHECK(map->full32.p, MH_ESTATE, "mh_map_get_full: map was built without keep_full");
    MH_HIP(hipSetDevice(map->ctx->device));
    const int nx = map->npnts[0], ny = map->npnts[1], nz = map->npnts[2];
    const long long nodes = (long long)nx * ny * nz;
    // the device keeps the nodes in the map's scan order (z fastest); the
    // caller receives config_makegrid order (x fastest)
    auto reorder = [&](auto *dst, const auto &src) {
        for (long long s = 0; s < nodes; s++) {
            const long long i = s / ((long long)ny * nz), j = (s / nz) % ny, k = s % nz;
            const long long m = i + (long long)nx * (j + (long long)ny * k);
            for (int c = 0; c < 8; c++) dst[8 * m + c] = src[8 * s + c];
        }
    };
    if (f32) {
        std::vector<float> h((size_t)(8 * nodes));
        MH_HIP(hipMemcpy(h.data(), map->full32.p, sizeof(float) * 8 * nodes, hipMemcpyDeviceToHost));
        reorder(f32, h);
    }
    if (f64) {
        MH_CHECK(map->full64.p, MH_ESTATE, "mh_map_get_full: no FP64 node values kept");
        std::vector<double> h((size_t)(8 * nodes));
        MH_HIP(hipMemcpy(h.data(), map->full64.p, sizeof(double) * 8 * nodes, hipMemcpyDeviceToHost));
        reorder(f64, h);
    }
    return MH_OK;
}

int mh_map_write(const mh_map *map, const char *path, int nzatoms, const char *names25,
                 const char *symbols2, const float *zparams7, float pcut_kJ) {
    MH_CHECK(map && path, MH_EINVAL, "mh_map_write: bad arguments");
    MH_CHECK(nzatoms >= 0 && (nzatoms == 0 || (names25 && symbols2 && zparams7)), MH_EINVAL,
             "mh_map_write: bad zatom arrays");
    const long long nodes = (long long)map->npnts[0] * map->npnts[1] * map->npnts[2];
    std::vector<int32_t> ptr(nodes);
    std::vector<float> v8(8 * (size_t)map->nfsize);
    int rc = mh_map_get(map, ptr.data(), v8.data());
    if (rc) return rc;
    FILE *f = fopen(path, "wb");
    MH_CHECK(f, MH_EIO, std::string("mh_map_write: cannot open ") + path);
    bool ok = true;
    char s40[40], s60[60], s25[25], s27[27];
    float rj = 0.0f;
    int32_t ij = 0, zt[11] = {0};
    padcpy(s40, "file contains header", 40);
    padcpy(s60, "Old type header 2003", 60);
    padcpy(s25, "probe", 25);
    ok = ok && rec_write(f, s40, 40) && rec_write(f, s60, 60) && rec_write(f, s25, 25) &&
         rec_write(f, &rj, 4);
    int32_t nz = nzatoms;
    ok = ok && rec_write(f, &nz, 4) && rec_write(f, zt, sizeof zt);
    for (int i = 0; i < nzatoms && ok; i++) {
        memcpy(s27, names25 + 25 * i, 25);
        memcpy(s27 + 25, symbols2 + 2 * i, 2);
        const float *zp = zparams7 + 7 * i;
        ok = rec_write(f, s27, 27) && rec_write(f, zp, 4) && rec_write(f, zp + 1, 16) &&
             rec_write(f, zp + 5, 8);
    }
    int32_t nb[3] = {map->npnts[0], map->npnts[1], map->npnts[2]};
    int32_t nf = (int32_t)map->nfsize;
    ok = ok && rec_write(f, &ij, 4) && rec_write(f, &pcut_kJ, 4) && rec_write(f, &rj, 4) &&
         rec_write(f, &rj, 4) && rec_write(f, &ij, 4) && rec_write(f, &rj, 4) &&
         rec_write(f, nb, 12) && rec_write(f, &nf, 4) && rec_write(f, ptr.data(), 4 * (size_t)nodes);
    for (int a = 0; a < 8 && ok; a++)
        ok = rec_write(f, v8.data() + (size_t)a * map->nfsize, 4 * (size_t)map->nfsize);
    if (fclose(f) != 0) ok = false;
    MH_CHECK(ok, MH_EIO, std::string("mh_map_write: write failed for ") + path);
    return MH_OK;
}

static int upload_map(mh_ctx *ctx, const int np[3], int64_t nf, const int32_t *ptr, const float *v8,
                      const double fund_ell[3], mh_map **out) {
    MH_CHECK(np[0] >= 2 && np[1] >= 2 && np[2] >= 2, MH_EINVAL, "map: bad grid size");
    MH_CHECK(nf >= 0 && nf < (1LL << 31), MH_EINVAL, "map: bad nfsize");
    const long long nodes = (long long)np[0] * np[1] * np[2];
    for (long long i = 0; i < nodes; i++)
        MH_CHECK(ptr[i] >= 0 && ptr[i] <= nf, MH_EINVAL, "map: ptr entry out of range");
    MH_HIP(hipSetDevice(ctx->device));
    mh_map *m = new mh_map();
    m->ctx = ctx;
    for (int i = 0; i < 3; i++) {
        m->npnts[i] = np[i];
        m->fund_ell[i] = fund_ell[i];
        m->step[i] = fund_ell[i] / ((double)np[i] - 1.0);  // mapheader.F90:140-146
    }
    // a context with a slanted cell: fund_ell is the map's effective box,
    // points fold into the context's slanted fundamental cell
    if (ctx->slanted) {
        m->slanted = true;
        for (int i = 0; i < 3; i++) m->fund_ell[i] = ctx->fund_ell[i];
        for (int i = 0; i < 9; i++) {
            m->U[i] = ctx->U[i];
            m->S[i] = ctx->S[i];
        }
    }
    m->nfsize = nf;
    m->counts[2] = nf;
    int rc;
    DBuf soa;
    if ((rc = m->ptr.alloc(sizeof(int32_t) * nodes)) || (rc = m->vals.alloc(sizeof(float) * 8 * (nf + 1))) ||
        (rc = soa.alloc(sizeof(float) * 8 * (nf + 1)))) {
        delete m;
        return rc;
    }
    if (hipMemcpy(m->ptr.p, ptr, sizeof(int32_t) * nodes, hipMemcpyHostToDevice) != hipSuccess ||
        (nf && hipMemcpy(soa.p, v8, sizeof(float) * 8 * nf, hipMemcpyHostToDevice) != hipSuccess)) {
        delete m;
        set_error("map upload failed");
        return MH_EHIP;
    }
    if (nf) {
        hipLaunchKernelGGL(k_soa_to_aos, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, ctx->stream,
                           (long long)nf, soa.as<float>(), m->vals.as<float>());
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) {
            delete m;
            set_error("map upload: transpose failed");
            return MH_EHIP;
        }
    }
    m->retained = true;
    ctx_retain(ctx);
    *out = m;
    return MH_OK;
}

int mh_map_from_host(mh_ctx *ctx, const int npnts[3], int64_t nfsize, const int32_t *ptr,
                     const float *vals8, const double fund_ell[3], mh_map **out) {
    MH_CHECK(ctx && npnts && ptr && (vals8 || nfsize == 0) && fund_ell && out, MH_EINVAL,
             "mh_map_from_host: bad arguments");
    return upload_map(ctx, npnts, nfsize, ptr, vals8, fund_ell, out);
}

int mh_map_load(mh_ctx *ctx, const char *path, const double fund_ell[3], mh_map **out) {
    MH_CHECK(ctx && path && fund_ell && out, MH_EINVAL, "mh_map_load: bad arguments");
    FILE *f = fopen(path, "rb");
    MH_CHECK(f, MH_EIO, std::string("mh_map_load: cannot open ") + path);
    char buf[128];
    size_t n = 0;
    bool ok = rec_read(f, buf, sizeof buf, &n) && n == 40 && strncmp(buf, "file contains header", 20) == 0;
    for (int r = 0; r < 3 && ok; r++) ok = rec_read(f, buf, sizeof buf, &n);
    int32_t nz = 0;
    ok = ok && rec_read(f, &nz, 4, &n) && nz >= 0 && nz < 100000 && rec_read(f, buf, sizeof buf, &n);
    for (int i = 0; i < nz && ok; i++)
        for (int r = 0; r < 4 && ok; r++) ok = rec_read(f, buf, sizeof buf, &n);
    for (int r = 0; r < 6 && ok; r++) ok = rec_read(f, buf, sizeof buf, &n);
    int32_t nb[3] = {0, 0, 0}, nf = 0;
    ok = ok && rec_read(f, nb, 12, &n) && n == 12 && rec_read(f, &nf, 4, &n);
    ok = ok && nb[0] > 1 && nb[1] > 1 && nb[2] > 1 && nf >= 0;
    std::vector<int32_t> ptr;
    std::vector<float> v8;
    if (ok) {
        long long nodes = (long long)nb[0] * nb[1] * nb[2];
        ptr.resize(nodes);
        v8.resize(8 * (size_t)nf);
        ok = rec_read(f, ptr.data(), 4 * (size_t)nodes, &n) && n == 4 * (size_t)nodes;
        for (int a = 0; a < 8 && ok; a++)
            ok = rec_read(f, v8.data() + (size_t)a * nf, 4 * (size_t)nf, &n) && n == 4 * (size_t)nf;
    }
    fclose(f);
    MH_CHECK(ok, MH_EIO, std::string("mh_map_load: not a map file or truncated: ") + path);
    int np[3] = {nb[0], nb[1], nb[2]};
    return upload_map(ctx, np, nf, ptr.data(), v8.data(), fund_ell, out);
}

int mh_map_interp(mh_ctx *ctx, const mh_map *map, int n, const double *xyz, double factor,
                  double *e, uint8_t *ok) {
    MH_CHECK(ctx && map && (n == 0 || (xyz && e && ok)), MH_EINVAL, "mh_map_interp: bad arguments");
    MH_CHECK(n >= 0, MH_EINVAL, "mh_map_interp: n < 0");
    if (n == 0) return MH_OK;
    MH_HIP(hipSetDevice(ctx->device));
    DBuf dx, de, dok;
    int rc;
    if ((rc = dx.alloc(sizeof(double) * 3 * n)) || (rc = de.alloc(sizeof(double) * n)) ||
        (rc = dok.alloc(n)))
        return rc;
    MH_HIP(hipMemcpyAsync(dx.p, xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream));
    SlantCell sc;
    for (int i = 0; i < 9; i++) {
        sc.U[i] = map->U[i];
        sc.S[i] = map->S[i];
    }
    hipLaunchKernelGGL(k_interp, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, dx.as<double>(),
                       map->ptr.as<int32_t>(), map->vals.as<float>(), map->npnts[0], map->npnts[1],
                       map->npnts[2], map->step[0], map->step[1], map->step[2], map->fund_ell[0],
                       map->fund_ell[1], map->fund_ell[2], factor, map->slanted ? 1 : 0, sc, de.as<double>(),
                       dok.as<uint8_t>());
    MH_HIP(hipGetLastError());
    MH_HIP(hipMemcpyAsync(e, de.p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipMemcpyAsync(ok, dok.p, n, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    return MH_OK;
}

}  // extern "C"
