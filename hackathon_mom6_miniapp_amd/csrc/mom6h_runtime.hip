// mom6h_runtime.hip -- runtime helpers, error reporting and the driver's
// inline loops (transports, velocity updates, periodic halo refill).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "mom6h_internal.h"

namespace mom6h {

static thread_local char g_err[1024] = "";

int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int dev_alloc(double **p, size_t count, bool zero)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(double));
    if (e != hipSuccess)
        return set_error(MOM6H_ENOMEM, "hipMalloc(%zu doubles): %s", count, hipGetErrorString(e));
    if (zero) MOM6H_CHECK(hipMemset(*p, 0, count * sizeof(double)));
    return MOM6H_OK;
}

int dev_upload(double **p, const double *host, size_t count)
{
    MOM6H_REQUIRE(host != nullptr, "missing host array");
    int rc = dev_alloc(p, count, false);
    if (rc) return rc;
    MOM6H_CHECK(hipMemcpy(*p, host, count * sizeof(double), hipMemcpyHostToDevice));
    return MOM6H_OK;
}

void dev_free(double *p)
{
    if (p) (void)hipFree(p);
}

// ---------------------------------------------------------------- halo fill
// One thread per (halo cell, k).  x-fill: columns isd..isc-1 and iec+1..ied of
// every row; y-fill: rows jsd..jsc-1 and jec+1..jed of every column, run after
// the x-fill so corners take their x-filled source.
template <typename T>
__device__ __forceinline__ void halo_x_body(const Dims &d, T *a, int nk, long long t)
{
    const int hw_w = d.isc - d.isd, hw_e = d.ied - d.iec, hw = hw_w + hw_e;
    const long long per_k = (long long)hw * d.n2;
    if (t >= per_k * nk) return;
    const int k = (int)(t / per_k) + 1;
    const int r = (int)(t % per_k);
    const int j = d.jsd + r / hw;
    const int c = r % hw;
    const int P = d.iec - d.isc + 1;
    const int i = (c < hw_w) ? d.isd + c : d.iec + 1 + (c - hw_w);
    const int src = (c < hw_w) ? i + P : i - P;
    a[IX3(d, i, j, k)] = a[IX3(d, src, j, k)];
}

template <typename T>
__global__ void k_halo_x(Dims d, T *a, int nk)
{
    halo_x_body(d, a, nk, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

template <typename T>
__device__ __forceinline__ void halo_y_body(const Dims &d, T *a, int nk, long long t)
{
    const int hs = d.jsc - d.jsd, hn = d.jed - d.jec, hw = hs + hn;
    const long long per_k = (long long)hw * d.n1;
    if (t >= per_k * nk) return;
    const int k = (int)(t / per_k) + 1;
    const int r = (int)(t % per_k);
    const int i = d.isd + r % d.n1;
    const int c = r / d.n1;
    const int Q = d.jec - d.jsc + 1;
    const int j = (c < hs) ? d.jsd + c : d.jec + 1 + (c - hs);
    const int src = (c < hs) ? j + Q : j - Q;
    a[IX3(d, i, j, k)] = a[IX3(d, i, src, k)];
}

template <typename T>
__global__ void k_halo_y(Dims d, T *a, int nk)
{
    halo_y_body(d, a, nk, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// the ensemble form: every (member, array) job of one direction in one launch
template <typename T, bool Y>
__global__ void __launch_bounds__(256) k_halo_batch(const HaloJob<T> *__restrict__ jobs, BatchMap mp)
{
    int m, local;
    batch_locate(mp, blockIdx.x, m, local);
    const HaloJob<T> &j = jobs[m];
    const long long t = (long long)local * 256 + threadIdx.x;
    if (Y) halo_y_body(j.d, j.a, j.nk, t);
    else halo_x_body(j.d, j.a, j.nk, t);
}

// jobs: device copy of h_jobs; x jobs are those with px, y jobs those with py
template <typename T>
int halo_batch(const HaloJob<T> *d_jobs, const HaloJob<T> *h_jobs, int n, hipStream_t s)
{
    MOM6H_REQUIRE(n <= kMaxMembers, "halo batch: at most %d jobs", kMaxMembers);
    for (int dir = 0; dir < 2; ++dir) {
        BatchMap mp{};
        mp.n = n;
        for (int m = 0; m < n; ++m) {
            const Dims &d = h_jobs[m].d;
            const bool on = dir == 0 ? h_jobs[m].px : h_jobs[m].py;
            const long long cells = dir == 0 ? (long long)((d.isc - d.isd) + (d.ied - d.iec)) * d.n2
                                             : (long long)((d.jsc - d.jsd) + (d.jed - d.jec)) * d.n1;
            mp.nbx[m] = on ? (int)((cells * h_jobs[m].nk + 255) / 256) : 0;
            mp.first[m + 1] = mp.first[m] + mp.nbx[m];
        }
        if (mp.first[n] == 0) continue;
        if (dir == 0) k_halo_batch<T, false><<<(unsigned)mp.first[n], 256, 0, s>>>(d_jobs, mp);
        else k_halo_batch<T, true><<<(unsigned)mp.first[n], 256, 0, s>>>(d_jobs, mp);
        MOM6H_LAUNCH_CHECK();
    }
    return MOM6H_OK;
}

template int halo_batch<float>(const HaloJob<float> *, const HaloJob<float> *, int, hipStream_t);
template int halo_batch<double>(const HaloJob<double> *, const HaloJob<double> *, int, hipStream_t);

template <typename T>
int halo_fill_t(const Dims &d, const Halo &hm, T *a, int nk, hipStream_t s)
{
    const int bs = 256;
    if (hm.px) {
        const long long n = (long long)((d.isc - d.isd) + (d.ied - d.iec)) * d.n2 * nk;
        if (n > 0) {
            k_halo_x<<<(unsigned)((n + bs - 1) / bs), bs, 0, s>>>(d, a, nk);
            MOM6H_LAUNCH_CHECK();
        }
    }
    if (hm.py) {
        const long long n = (long long)((d.jsc - d.jsd) + (d.jed - d.jec)) * d.n1 * nk;
        if (n > 0) {
            k_halo_y<<<(unsigned)((n + bs - 1) / bs), bs, 0, s>>>(d, a, nk);
            MOM6H_LAUNCH_CHECK();
        }
    }
    return MOM6H_OK;
}

int halo_fill(const Dims &d, const Halo &hm, double *a, int nk, hipStream_t s)
{
    return halo_fill_t(d, hm, a, nk, s);
}

int halo_fill(const Dims &d, const Halo &hm, float *a, int nk, hipStream_t s)
{
    return halo_fill_t(d, hm, a, nk, s);
}

// ------------------------------------------------------------ conversions
__global__ void k_narrow(const double *__restrict__ src, float *__restrict__ dst, long long n)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = (float)src[t];
}

__global__ void k_widen(const float *__restrict__ src, double *__restrict__ dst, long long n)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = (double)src[t];
}

int narrow_to_f32(const double *src, float *dst, long long n, hipStream_t s)
{
    if (n <= 0) return MOM6H_OK;
    k_narrow<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(src, dst, n);
    MOM6H_LAUNCH_CHECK();
    return MOM6H_OK;
}

int widen_to_f64(const float *src, double *dst, long long n, hipStream_t s)
{
    if (n <= 0) return MOM6H_OK;
    k_widen<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(src, dst, n);
    MOM6H_LAUNCH_CHECK();
    return MOM6H_OK;
}

int dev_alloc_bytes(void **p, size_t bytes, bool zero)
{
    *p = nullptr;
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return set_error(MOM6H_ENOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    if (zero) MOM6H_CHECK(hipMemset(*p, 0, bytes));
    return MOM6H_OK;
}

// --------------------------------------------------------------- transports
// compute_transports, app/rk2_serial_driver.F90:544-561:
//   uh = u*0.5*(h + h(min(i+1,ied)))*dyCu ; vh = v*0.5*(h + h(i,min(j+1,jed)))*dxCv
// A block owns 64 x 8 columns (64 x 4 threads, two rows each) and walks a
// range of layers: the 2D metrics of its columns are loaded once, and the
// h(i, j+1) row it reads is the next thread row's own, so it comes from the
// CU's cache instead of a second fetch by a block on another XCD.
constexpr int TR_X = 64, TR_Y = 8;

__global__ void __launch_bounds__(256) k_transports(Dims d, const double *__restrict__ u,
                                                    const double *__restrict__ v, const double *__restrict__ h,
                                                    const double *__restrict__ dyCu,
                                                    const double *__restrict__ dxCv, double *__restrict__ uh,
                                                    double *__restrict__ vh)
{
    const int nxy = gridDim.x * gridDim.y;
    const int tile = xcd_tile(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), nxy * gridDim.z);
    const int xy = tile % nxy, kz = tile / nxy;
    const int per = (d.nk + (int)gridDim.z - 1) / (int)gridDim.z;
    const int k0 = 1 + kz * per, k1 = k0 + per - 1 < d.nk ? k0 + per - 1 : d.nk;
    const int i = d.isd + (xy % (int)gridDim.x) * TR_X + (int)(threadIdx.x % TR_X);
    const int jt = d.jsd + (xy / (int)gridDim.x) * TR_Y + (int)(threadIdx.x / TR_X);
    if (i > d.ied) return;
    const int ip = i + 1 < d.ied ? i + 1 : d.ied;
    double my[2] = {0.0, 0.0}, mx[2] = {0.0, 0.0};
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int j = jt + 4 * r;
        if (j <= d.jed) {
            my[r] = dyCu[IX2(d, i, j)];
            mx[r] = dxCv[IX2(d, i, j)];
        }
    }
    for (int k = k0; k <= k1; ++k) {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int j = jt + 4 * r;
            if (j > d.jed) continue;
            const int jp = j + 1 < d.jed ? j + 1 : d.jed;
            const long long c = IX3(d, i, j, k);
            const double hc = h[c];
            uh[c] = u[c] * 0.5 * (hc + h[IX3(d, ip, j, k)]) * my[r];
            vh[c] = v[c] * 0.5 * (hc + h[IX3(d, i, jp, k)]) * mx[r];
        }
    }
}

// ---------------------------------------------------------- velocity update
// app/rk2_serial_driver.F90:257-270 (fac_dt = dt) and :337-350 (0.5*dt):
// ou(i,j,k) = bu + fac_dt*(CAu + diffu) for i in isc..iec-1, j in jsc..jec;
// ov likewise for i in isc..iec, j in jsc..jec-1.
__global__ void k_vel_update(Dims d, double fdt, const double *bu, const double *bv,
                             const double *__restrict__ CAu, const double *__restrict__ diffu,
                             const double *__restrict__ CAv, const double *__restrict__ diffv,
                             double *ou, double *ov)
{
    const int i = d.isc + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.jsc + blockIdx.y;
    const int k = 1 + blockIdx.z;
    if (i > d.iec) return;
    const long long c = IX3(d, i, j, k);
    if (i <= d.iec - 1) ou[c] = bu[c] + fdt * (CAu[c] + diffu[c]);
    if (j <= d.jec - 1) ov[c] = bv[c] + fdt * (CAv[c] + diffv[c]);
}


// ------------------------------------------------------------ seeded state
// Counter-based N(mean, sigma) samples: element e of field f sums 12 uniforms
// U_r = (x_r >> 11) * 2^-53, x_r = SplitMix64(seed*2^32 + f*2^52 + 12e + r),
// left to right from 0, then sigma*(sum - 6) [+ mean] [* mask2d].  Integer
// arithmetic, exact conversions and IEEE adds: the host generator
// (hackathon_mom6_miniapp_amd/state.py) produces the same bits.
__global__ void k_seeded_normal(double *out, long long n, unsigned long long seed,
                                unsigned long long field, int add_mean, double mean, double sigma,
                                const double *mask2d, long long n2)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const unsigned long long base = (seed << 32) + (field << 52) + (unsigned long long)e * 12ull;
    double acc = 0.0;
    for (int r = 0; r < 12; ++r) {
        unsigned long long z = base + (unsigned long long)r;
        z += 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z = z ^ (z >> 31);
        acc = acc + (double)(z >> 11) * 0x1p-53;
    }
    double v = sigma * (acc - 6.0);
    if (add_mean) v = mean + v;
    if (mask2d) v = v * mask2d[e % n2];
    out[e] = v;
}

}  // namespace mom6h

using namespace mom6h;

extern "C" {

const char *mom6h_last_error(void) { return g_err; }
const char *mom6h_version(void) { return "mom6hip 0.1 (gfx950)"; }

int mom6h_set_device(int device)
{
    MOM6H_CHECK(hipSetDevice(device));
    return MOM6H_OK;
}
int mom6h_device_synchronize(void)
{
    MOM6H_CHECK(hipDeviceSynchronize());
    return MOM6H_OK;
}
int mom6h_malloc(void **p, size_t bytes)
{
    MOM6H_REQUIRE(p != nullptr, "mom6h_malloc: NULL out pointer");
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess)
        return set_error(MOM6H_ENOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    return MOM6H_OK;
}
int mom6h_free(void *p)
{
    if (p) MOM6H_CHECK(hipFree(p));
    return MOM6H_OK;
}
int mom6h_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    MOM6H_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return MOM6H_OK;
}
int mom6h_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    MOM6H_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return MOM6H_OK;
}
int mom6h_memcpy_d2d(void *dst, const void *src, size_t bytes)
{
    MOM6H_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return MOM6H_OK;
}
int mom6h_memset(void *p, int value, size_t bytes)
{
    MOM6H_CHECK(hipMemset(p, value, bytes));
    return MOM6H_OK;
}
int mom6h_stream_create(void **stream)
{
    MOM6H_REQUIRE(stream != nullptr, "mom6h_stream_create: NULL out pointer");
    hipStream_t s;
    MOM6H_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void *)s;
    return MOM6H_OK;
}
int mom6h_stream_destroy(void *stream)
{
    if (stream) MOM6H_CHECK(hipStreamDestroy((hipStream_t)stream));
    return MOM6H_OK;
}
int mom6h_stream_synchronize(void *stream)
{
    MOM6H_CHECK(hipStreamSynchronize((hipStream_t)stream));
    return MOM6H_OK;
}

int mom6h_transports(int isd, int ied, int jsd, int jed, int nk, const double *u, const double *v,
                     const double *h, const double *dyCu, const double *dxCv, double *uh,
                     double *vh, void *stream)
{
    MOM6H_REQUIRE(ied >= isd && jed >= jsd && nk >= 1, "mom6h_transports: bad bounds");
    MOM6H_REQUIRE(u && v && h && dyCu && dxCv && uh && vh, "mom6h_transports: NULL array");
    Dims d = make_dims(isd, ied, jsd, jed, isd, ied, jsd, jed, nk);
    const dim3 g((unsigned)((d.n1 + TR_X - 1) / TR_X), (unsigned)((d.n2 + TR_Y - 1) / TR_Y),
                 (unsigned)((nk + 14) / 15));   // about 15 layers per block
    k_transports<<<g, 256, 0, (hipStream_t)stream>>>(d, u, v, h, dyCu, dxCv, uh, vh);
    MOM6H_LAUNCH_CHECK();
    return MOM6H_OK;
}

int mom6h_vel_update(int isd, int ied, int jsd, int jed, int isc, int iec, int jsc, int jec, int nk,
                     double fac_dt, const double *bu, const double *bv, const double *CAu,
                     const double *diffu, const double *CAv, const double *diffv, double *ou,
                     double *ov, void *stream)
{
    MOM6H_REQUIRE(isd <= isc && isc <= iec && iec <= ied && jsd <= jsc && jsc <= jec && jec <= jed && nk >= 1,
                  "mom6h_vel_update: bad bounds");
    MOM6H_REQUIRE(bu && bv && CAu && diffu && CAv && diffv && ou && ov, "mom6h_vel_update: NULL array");
    Dims d = make_dims(isd, ied, jsd, jed, isc, iec, jsc, jec, nk);
    const int ni = iec - isc + 1, nj = jec - jsc + 1;
    dim3 b(128), g((ni + 127) / 128, nj, nk);
    k_vel_update<<<g, b, 0, (hipStream_t)stream>>>(d, fac_dt, bu, bv, CAu, diffu, CAv, diffv, ou, ov);
    MOM6H_LAUNCH_CHECK();
    return MOM6H_OK;
}

int mom6h_halo_fill(int isd, int ied, int jsd, int jed, int isc, int iec, int jsc, int jec, int nk,
                    int periodic_x, int periodic_y, double *a, void *stream)
{
    MOM6H_REQUIRE(a != nullptr, "mom6h_halo_fill: NULL array");
    MOM6H_REQUIRE(isd <= isc && isc <= iec && iec <= ied && jsd <= jsc && jsc <= jec && jec <= jed && nk >= 1,
                  "mom6h_halo_fill: bad bounds");
    MOM6H_REQUIRE(!periodic_x || (iec - isc + 1) >= (isc - isd) && (iec - isc + 1) >= (ied - iec),
                  "mom6h_halo_fill: compute domain narrower than the x halo");
    MOM6H_REQUIRE(!periodic_y || (jec - jsc + 1) >= (jsc - jsd) && (jec - jsc + 1) >= (jed - jec),
                  "mom6h_halo_fill: compute domain narrower than the y halo");
    Dims d = make_dims(isd, ied, jsd, jed, isc, iec, jsc, jec, nk);
    Halo hm{periodic_x, periodic_y};
    return halo_fill(d, hm, a, nk, (hipStream_t)stream);
}

int mom6h_bt_cfl(double dt, int nsteps, double dx_km, double depth, double *cfl, double *cfl_2d, int *min_nsteps,
                 int *status)
{
    MOM6H_REQUIRE(cfl && cfl_2d && min_nsteps && status, "mom6h_bt_cfl: NULL output");
    MOM6H_REQUIRE(dt > 0.0 && nsteps >= 1 && dx_km > 0.0 && depth > 0.0, "mom6h_bt_cfl: bad arguments");
    // check_bt_cfl (app/rk2_serial_driver.F90:625-662), G_EARTH = 9.80 (mom6_types.F90:80)
    const double dx_m = dx_km * 1000.0;
    const double dtbt = dt / (double)nsteps;
    const double c_grav = sqrt(9.80 * depth);
    *cfl = c_grav * dtbt / dx_m;
    *cfl_2d = *cfl * sqrt(2.0);
    *min_nsteps = (int)ceil(dt * c_grav / dx_m * 2.5);
    *status = *cfl_2d > 0.9 ? 2 : (*cfl_2d > 0.7 ? 1 : 0);
    return MOM6H_OK;
}

int mom6h_seeded_normal(double *out, long count, unsigned long long seed, int field, int add_mean,
                        double mean, double sigma, const double *mask2d, long n2, void *stream)
{
    MOM6H_REQUIRE(out && count > 0 && field >= 0 && field < 4096, "mom6h_seeded_normal: bad arguments");
    MOM6H_REQUIRE(!mask2d || (n2 > 0 && count % n2 == 0), "mom6h_seeded_normal: count must be a multiple of n2");
    const int bs = 256;
    k_seeded_normal<<<(unsigned)((count + bs - 1) / bs), bs, 0, (hipStream_t)stream>>>(
        out, count, seed, (unsigned long long)field, add_mean, mean, sigma, mask2d, n2 > 0 ? n2 : 1);
    MOM6H_LAUNCH_CHECK();
    return MOM6H_OK;
}

}  // extern "C"
