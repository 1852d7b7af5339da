// mom6h_diag.hip -- the mini-app's diagnostics on the device.
//
// The fields of src/common/mom6_diag.F90 and app/rk2_serial_driver.F90 as the
// serial driver posts them (diag_mode = 1):
//   post_product_sum_u/v (mom6_diag.F90:276-340): vertical sums of a product
//     of two 3D fields on the u range (isc..iec-1, jsc..jec) or the v range
//     (isc..iec, jsc..jec-1) of a 2D work array, in k order per point;
//   compute_and_post_KE (rk2_serial_driver.F90:664-693): KE at h points;
//   compute_and_post_mass (:695-723): the column sum of h, in k order;
// and the statistics post_data_2d/3d print (mom6_diag.F90:166-274): min, max,
// mean and rms over the compute domain (all layers of a 3D field).
//
// Per-point fields are bit-identical to the serial loops.  The mean and the
// rms are reductions: the serial loops add the values one by one in (k, j, i)
// order; here every thread accumulates its values in index order as
// double-double sums (TwoSum; squares split exactly with an FMA), blocks and
// then one block combine them in a fixed order, and the sum is rounded once at
// the end.  The result is deterministic and at least as accurate as the
// serial sum; the two differ by the serial sum's own rounding (at most about
// count * 1.1e-16 of the sum of |values|).
#include "mom6h_internal.h"

namespace mom6h {

constexpr int kStatsBlocks = 256;    // partial results of the first reduction stage
constexpr int kStatsThreads = 256;

struct DD {   // unevaluated sum hi + lo
    double hi, lo;
};

__device__ inline DD dd_add(DD a, double b)
{
    const double s = a.hi + b;
    const double bb = s - a.hi;
    const double err = (a.hi - (s - bb)) + (b - bb);
    const double lo = a.lo + err;
    const double hi = s + lo;
    return DD{hi, lo - (hi - s)};
}

__device__ inline DD dd_add(DD a, DD b)
{
    DD r = dd_add(a, b.hi);
    return dd_add(r, b.lo);
}

__device__ inline DD dd_add_square(DD a, double x)
{
    const double p = x * x;
    const double e = fma(x, x, -p);   // x*x = p + e exactly
    return dd_add(dd_add(a, p), e);
}

struct Stats {
    double mn, mx;
    DD s, q;
};

__device__ inline Stats stats_merge(const Stats &a, const Stats &b)
{
    return Stats{fmin(a.mn, b.mn), fmax(a.mx, b.mx), dd_add(a.s, b.s), dd_add(a.q, b.q)};
}

// work_2d(i, j) = sum_k a(i,j,k)*b(i,j,k) on the u or v range (the reference
// zeroes the range, then adds the layers in order)
__global__ void k_diag_product_sum(Dims d, const double *__restrict__ a, const double *__restrict__ b, int i_hi,
                                   int j_hi, double *__restrict__ work)
{
    const int i = d.isc + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = d.jsc + (int)blockIdx.y;
    if (i > i_hi || j > j_hi) return;
    double acc = 0.0;
    for (int k = 1; k <= d.nk; ++k) {
        const long long c = IX3(d, i, j, k);
        acc = acc + a[c] * b[c];
    }
    work[IX2(d, i, j)] = acc;
}

// KE = 0.5*(0.5*(u(i)**2 + u(i-1)**2) + 0.5*(v(j)**2 + v(j-1)**2)) at the compute h points
__global__ void k_diag_ke(Dims d, const double *__restrict__ u, const double *__restrict__ v,
                          double *__restrict__ ke)
{
    const int i = d.isc + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = d.jsc + (int)blockIdx.y;
    const int k = 1 + (int)blockIdx.z;
    if (i > d.iec) return;
    const double u0 = u[IX3(d, i, j, k)], u1 = u[IX3(d, i - 1, j, k)];
    const double v0 = v[IX3(d, i, j, k)], v1 = v[IX3(d, i, j - 1, k)];
    ke[IX3(d, i, j, k)] = 0.5 * (0.5 * (u0 * u0 + u1 * u1) + 0.5 * (v0 * v0 + v1 * v1));
}

// mass = 0 over the data domain, then the column sums of h on the compute points
__global__ void k_diag_mass(Dims d, const double *__restrict__ h, double *__restrict__ mass)
{
    const int i = d.isd + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = d.jsd + (int)blockIdx.y;
    if (i > d.ied) return;
    double acc = 0.0;
    if (i >= d.isc && i <= d.iec && j >= d.jsc && j <= d.jec)
        for (int k = 1; k <= d.nk; ++k) acc = acc + h[IX3(d, i, j, k)];
    mass[IX2(d, i, j)] = acc;
}

__device__ inline void block_reduce(Stats &st, Stats *sh)
{
    const int t = threadIdx.x;
    sh[t] = st;
    __syncthreads();
    for (int w = kStatsThreads / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] = stats_merge(sh[t], sh[t + w]);
        __syncthreads();
    }
    st = sh[0];
}

// First stage: block b takes the elements b*T + t, b*T + t + B*T, ... of the
// compute domain (index = ((k-1)*nj + jj)*ni + ii), each thread in index order.
__global__ void __launch_bounds__(kStatsThreads) k_diag_stats_partial(Dims d, int nk, const double *__restrict__ f,
                                                                      double *__restrict__ part)
{
    __shared__ Stats sh[kStatsThreads];
    const long long ni = d.iec - d.isc + 1, nj = d.jec - d.jsc + 1, n = ni * nj * nk;
    Stats st{__builtin_huge_val(), -__builtin_huge_val(), DD{0.0, 0.0}, DD{0.0, 0.0}};
    for (long long e = (long long)blockIdx.x * kStatsThreads + threadIdx.x; e < n;
         e += (long long)kStatsBlocks * kStatsThreads) {
        const int ii = (int)(e % ni);
        const long long r = e / ni;
        const int jj = (int)(r % nj), k = (int)(r / nj) + 1;
        const double x = f[IX3(d, d.isc + ii, d.jsc + jj, k)];
        st.mn = fmin(st.mn, x);
        st.mx = fmax(st.mx, x);
        st.s = dd_add(st.s, x);
        st.q = dd_add_square(st.q, x);
    }
    block_reduce(st, sh);
    if (threadIdx.x == 0) {
        double *p = part + 6 * blockIdx.x;
        p[0] = st.mn; p[1] = st.mx; p[2] = st.s.hi; p[3] = st.s.lo; p[4] = st.q.hi; p[5] = st.q.lo;
    }
}

// Second stage: one block combines the partials; out = min, max, mean, rms
__global__ void __launch_bounds__(kStatsThreads) k_diag_stats_final(Dims d, int nk, const double *__restrict__ part,
                                                                    double *__restrict__ out)
{
    __shared__ Stats sh[kStatsThreads];
    Stats st{__builtin_huge_val(), -__builtin_huge_val(), DD{0.0, 0.0}, DD{0.0, 0.0}};
    for (int b = threadIdx.x; b < kStatsBlocks; b += kStatsThreads) {
        const double *p = part + 6 * b;
        st = stats_merge(st, Stats{p[0], p[1], DD{p[2], p[3]}, DD{p[4], p[5]}});
    }
    block_reduce(st, sh);
    if (threadIdx.x == 0) {
        const double cnt = (double)((long long)(d.iec - d.isc + 1) * (d.jec - d.jsc + 1) * nk);
        out[0] = st.mn;
        out[1] = st.mx;
        out[2] = (st.s.hi + st.s.lo) / cnt;
        out[3] = sqrt((st.q.hi + st.q.lo) / cnt);
    }
}

}  // namespace mom6h

using namespace mom6h;

extern "C" {

int mom6h_diag_product_sum(int isd, int ied, int jsd, int jed, int isc, int iec, int jsc, int jec, int nk,
                           const double *a, const double *b, int position, double *work_2d, void *stream)
{
    MOM6H_REQUIRE(a && b && work_2d, "mom6h_diag_product_sum: NULL array");
    MOM6H_REQUIRE(position == 0 || position == 1, "mom6h_diag_product_sum: position is 0 (u) or 1 (v)");
    MOM6H_REQUIRE(isd <= isc - 1 && iec <= ied && jsd <= jsc - 1 && jec <= jed && isc <= iec && jsc <= jec && nk >= 1,
                  "mom6h_diag_product_sum: bad bounds");
    const Dims d = make_dims(isd, ied, jsd, jed, isc, iec, jsc, jec, nk);
    const int i_hi = position == 0 ? iec - 1 : iec, j_hi = position == 0 ? jec : jec - 1;
    if (i_hi < isc || j_hi < jsc) return MOM6H_OK;
    const dim3 bs(64), g((unsigned)((i_hi - isc + 1 + 63) / 64), (unsigned)(j_hi - jsc + 1));
    k_diag_product_sum<<<g, bs, 0, (hipStream_t)stream>>>(d, a, b, i_hi, j_hi, work_2d);
    MOM6H_LAUNCH_CHECK();
    return MOM6H_OK;
}

int mom6h_diag_ke(int isd, int ied, int jsd, int jed, int isc, int iec, int jsc, int jec, int nk, const double *u,
                  const double *v, double *ke, void *stream)
{
    MOM6H_REQUIRE(u && v && ke, "mom6h_diag_ke: NULL array");
    MOM6H_REQUIRE(isd <= isc - 1 && iec <= ied && jsd <= jsc - 1 && jec <= jed && isc <= iec && jsc <= jec && nk >= 1,
                  "mom6h_diag_ke: bad bounds");
    const Dims d = make_dims(isd, ied, jsd, jed, isc, iec, jsc, jec, nk);
    const dim3 bs(64), g((unsigned)((iec - isc + 1 + 63) / 64), (unsigned)(jec - jsc + 1), (unsigned)nk);
    k_diag_ke<<<g, bs, 0, (hipStream_t)stream>>>(d, u, v, ke);
    MOM6H_LAUNCH_CHECK();
    return MOM6H_OK;
}

int mom6h_diag_mass(int isd, int ied, int jsd, int jed, int isc, int iec, int jsc, int jec, int nk, const double *h,
                    double *mass, void *stream)
{
    MOM6H_REQUIRE(h && mass, "mom6h_diag_mass: NULL array");
    MOM6H_REQUIRE(isd <= isc && iec <= ied && jsd <= jsc && jec <= jed && isc <= iec && jsc <= jec && nk >= 1,
                  "mom6h_diag_mass: bad bounds");
    const Dims d = make_dims(isd, ied, jsd, jed, isc, iec, jsc, jec, nk);
    const dim3 bs(64), g((unsigned)((d.n1 + 63) / 64), (unsigned)d.n2);
    k_diag_mass<<<g, bs, 0, (hipStream_t)stream>>>(d, h, mass);
    MOM6H_LAUNCH_CHECK();
    return MOM6H_OK;
}

long mom6h_diag_stats_work_doubles(void) { return 6L * kStatsBlocks; }

int mom6h_diag_stats(int isd, int ied, int jsd, int jed, int isc, int iec, int jsc, int jec, int nk,
                     const double *field, double *d_work, double *d_out, void *stream)
{
    MOM6H_REQUIRE(field && d_work && d_out, "mom6h_diag_stats: NULL array");
    MOM6H_REQUIRE(isd <= isc && iec <= ied && jsd <= jsc && jec <= jed && isc <= iec && jsc <= jec && nk >= 1,
                  "mom6h_diag_stats: bad bounds");
    const Dims d = make_dims(isd, ied, jsd, jed, isc, iec, jsc, jec, nk);
    hipStream_t s = (hipStream_t)stream;
    k_diag_stats_partial<<<kStatsBlocks, kStatsThreads, 0, s>>>(d, nk, field, d_work);
    MOM6H_LAUNCH_CHECK();
    k_diag_stats_final<<<1, kStatsThreads, 0, s>>>(d, nk, d_work, d_out);
    MOM6H_LAUNCH_CHECK();
    return MOM6H_OK;
}

}  // extern "C"
