// k_fft.hip -- the PME convolution (forward 3-D real FFT, multiplication by
// the reciprocal kernel G, inverse FFT) for grids whose planes fit in LDS,
// in three launches instead of six library transforms and a separate
// multiplication:
//   k_fft_planes_fwd  one workgroup per z-plane and grid set: real-to-complex
//                     DFT along x, then complex DFT along y, in LDS
//   k_fft_columns     one workgroup per y-row of the half-complex grid: DFT
//                     along z, multiplication by G, inverse DFT along z
//   k_fft_planes_inv  one workgroup per z-plane: inverse DFT along y, then
//                     complex-to-real along x
// Every transform is a direct DFT sum (each output a length-K dot product
// with a twiddle table held in LDS); at the grid sizes this path takes the
// work is small next to the launches and global round trips it replaces.
// Each lane computes R outputs of one frequency, so the input values it reads
// are shared by the lanes of other frequencies (LDS broadcasts) and only the
// twiddle differs per lane.
//
// Conventions are those of the library path it replaces: forward
// e^{-2 pi i jk/K}, inverse e^{+2 pi i jk/K}, both unnormalized; the real
// grid is (z*K1 + y)*K0 + x, the complex half grid (z*K1 + y)*H0 + m with
// H0 = K0/2 + 1; the complex-to-real step takes the real parts of the m = 0
// and m = K0/2 terms.  Sums run in a fixed order: bit-reproducible.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pmq_internal.h"

namespace pmq {

template <class T>
struct alignas(2 * sizeof(T)) cpx {
    T re, im;
};

constexpr int FFT_R = 4;            // outputs per lane
constexpr int FFT_MAX_LDS = 65536;  // bytes of LDS a workgroup may take

template <class T>
struct FftArgs {
    int K0, K1, K2, H0;
    const cpx<T> *tw; // forward twiddles e^{-2 pi i j/K} of axes 0, 1, 2 (K0 + K1 + K2 entries)
    T *real;          // nset real grids, nreal apart
    cpx<T> *cplx;     // nset complex half grids, ncplx apart
    const T *G;       // H0 K1 K2 reciprocal kernel
    size_t nreal, ncplx;
};

template <class T>
__device__ __forceinline__ void cmad(cpx<T> &acc, const cpx<T> &a, const cpx<T> &w)
{
    acc.re += a.re * w.re - a.im * w.im;
    acc.im += a.re * w.im + a.im * w.re;
}

// acc += a * conj(w)
template <class T>
__device__ __forceinline__ void cmadc(cpx<T> &acc, const cpx<T> &a, const cpx<T> &w)
{
    acc.re += a.re * w.re + a.im * w.im;
    acc.im += a.im * w.re - a.re * w.im;
}

// Complex DFT along the slow index of an L x W array in LDS (row j holds
// W values, w[j*W + m]), for all W columns: out(k, m) = sum_j w[j*W + m]
// tw[(k j) mod L] (conjugated twiddles for the inverse).  Task t: frequency
// k = t % L, columns m0 .. m0+R-1 with m0 = (t / L) R.  The results stay in
// registers (acc) until the caller has synchronized.
template <class T, int NT, bool INV>
__device__ __forceinline__ void dft_slow(const cpx<T> *buf, const cpx<T> *tw, int L, int W, int tid, int bd,
                                         cpx<T> (&acc)[NT][FFT_R])
{
    const int ngm = (W + FFT_R - 1) / FFT_R;
#pragma unroll
    for (int q = 0; q < NT; q++) {
#pragma unroll
        for (int r = 0; r < FFT_R; r++) acc[q][r] = cpx<T>{T(0), T(0)};
        const int t = tid + q * bd;
        if (t >= L * ngm) continue;
        const int k = t % L, m0 = (t / L) * FFT_R;
        int col[FFT_R];
#pragma unroll
        for (int r = 0; r < FFT_R; r++) col[r] = min(m0 + r, W - 1);
        int idx = 0;
        for (int j = 0; j < L; j++) {
            const cpx<T> w = tw[idx];
            idx += k;
            if (idx >= L) idx -= L;
            const cpx<T> *row = buf + (size_t)j * W;
#pragma unroll
            for (int r = 0; r < FFT_R; r++) {
                if (INV) cmadc(acc[q][r], row[col[r]], w);
                else cmad(acc[q][r], row[col[r]], w);
            }
        }
    }
}

// write the results of dft_slow back as row k (G: optional real factor
// table of the same L x W layout, rows `gstride` apart, offset goff)
template <class T, int NT>
__device__ __forceinline__ void dft_slow_store(cpx<T> *buf, int L, int W, int tid, int bd,
                                               const cpx<T> (&acc)[NT][FFT_R], const T *G = nullptr,
                                               size_t gstride = 0, size_t goff = 0)
{
    const int ngm = (W + FFT_R - 1) / FFT_R;
#pragma unroll
    for (int q = 0; q < NT; q++) {
        const int t = tid + q * bd;
        if (t >= L * ngm) continue;
        const int k = t % L, m0 = (t / L) * FFT_R;
#pragma unroll
        for (int r = 0; r < FFT_R; r++) {
            const int m = m0 + r;
            if (m >= W) continue;
            cpx<T> v = acc[q][r];
            if (G) {
                const T g = G[(size_t)k * gstride + goff + m];
                v.re *= g;
                v.im *= g;
            }
            buf[(size_t)k * W + m] = v;
        }
    }
}

template <class T, int NT>
__global__ __launch_bounds__(1024) void k_fft_planes_fwd(FftArgs<T> a)
{
    extern __shared__ double fft_lds[];
    const int K0 = a.K0, K1 = a.K1, H0 = a.H0;
    const int z = blockIdx.x, set = blockIdx.y, tid = threadIdx.x, bd = blockDim.x;
    cpx<T> *buf = reinterpret_cast<cpx<T> *>(fft_lds);
    cpx<T> *w0 = buf + (size_t)H0 * K1, *w1 = w0 + K0;
    T *rb = reinterpret_cast<T *>(buf); // the real plane, K0 x K1, in the same space
    const T *in = a.real + set * a.nreal + (size_t)z * K0 * K1;
    for (int i = tid; i < K0 * K1; i += bd) rb[i] = in[i];
    for (int i = tid; i < K0 + K1; i += bd) w0[i] = a.tw[i];
    __syncthreads();
    // x: real to half-complex; task = frequency m, rows y0 .. y0+R-1
    cpx<T> acc[NT][FFT_R];
    const int ngy = (K1 + FFT_R - 1) / FFT_R;
#pragma unroll
    for (int q = 0; q < NT; q++) {
#pragma unroll
        for (int r = 0; r < FFT_R; r++) acc[q][r] = cpx<T>{T(0), T(0)};
        const int t = tid + q * bd;
        if (t >= H0 * ngy) continue;
        const int m = t % H0, y0 = (t / H0) * FFT_R;
        const T *row[FFT_R];
#pragma unroll
        for (int r = 0; r < FFT_R; r++) row[r] = rb + (size_t)min(y0 + r, K1 - 1) * K0;
        int idx = 0;
        for (int x = 0; x < K0; x++) {
            const cpx<T> w = w0[idx];
            idx += m;
            if (idx >= K0) idx -= K0;
#pragma unroll
            for (int r = 0; r < FFT_R; r++) {
                const T d = row[r][x];
                acc[q][r].re += d * w.re;
                acc[q][r].im += d * w.im;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NT; q++) {
        const int t = tid + q * bd;
        if (t >= H0 * ngy) continue;
        const int m = t % H0, y0 = (t / H0) * FFT_R;
#pragma unroll
        for (int r = 0; r < FFT_R; r++)
            if (y0 + r < K1) buf[(size_t)(y0 + r) * H0 + m] = acc[q][r];
    }
    __syncthreads();
    // y: complex DFT along y of the H0 columns
    dft_slow<T, NT, false>(buf, w1, K1, H0, tid, bd, acc);
    __syncthreads();
    dft_slow_store<T, NT>(buf, K1, H0, tid, bd, acc);
    __syncthreads();
    cpx<T> *out = a.cplx + set * a.ncplx + (size_t)z * K1 * H0;
    for (int i = tid; i < K1 * H0; i += bd) out[i] = buf[i];
}

template <class T, int NT>
__global__ __launch_bounds__(1024) void k_fft_columns(FftArgs<T> a)
{
    extern __shared__ double fft_lds[];
    const int K1 = a.K1, K2 = a.K2, H0 = a.H0;
    const int y = blockIdx.x, set = blockIdx.y, tid = threadIdx.x, bd = blockDim.x;
    cpx<T> *buf = reinterpret_cast<cpx<T> *>(fft_lds); // K2 x H0: row z
    cpx<T> *w2 = buf + (size_t)K2 * H0;
    cpx<T> *cg = a.cplx + set * a.ncplx;
    for (int i = tid; i < K2 * H0; i += bd) {
        const int z = i / H0, m = i - z * H0;
        buf[i] = cg[((size_t)z * K1 + y) * H0 + m];
    }
    for (int i = tid; i < K2; i += bd) w2[i] = a.tw[a.K0 + K1 + i];
    __syncthreads();
    cpx<T> acc[NT][FFT_R];
    dft_slow<T, NT, false>(buf, w2, K2, H0, tid, bd, acc);
    __syncthreads();
    // times G(m, y, k3), G rows (k3) K1 H0 apart
    dft_slow_store<T, NT>(buf, K2, H0, tid, bd, acc, a.G, (size_t)K1 * H0, (size_t)y * H0);
    __syncthreads();
    dft_slow<T, NT, true>(buf, w2, K2, H0, tid, bd, acc);
    __syncthreads();
    dft_slow_store<T, NT>(buf, K2, H0, tid, bd, acc);
    __syncthreads();
    for (int i = tid; i < K2 * H0; i += bd) {
        const int z = i / H0, m = i - z * H0;
        cg[((size_t)z * K1 + y) * H0 + m] = buf[i];
    }
}

template <class T, int NT>
__global__ __launch_bounds__(1024) void k_fft_planes_inv(FftArgs<T> a)
{
    extern __shared__ double fft_lds[];
    const int K0 = a.K0, K1 = a.K1, H0 = a.H0;
    const int z = blockIdx.x, set = blockIdx.y, tid = threadIdx.x, bd = blockDim.x;
    cpx<T> *buf = reinterpret_cast<cpx<T> *>(fft_lds);
    cpx<T> *w0 = buf + (size_t)H0 * K1, *w1 = w0 + K0;
    const cpx<T> *in = a.cplx + set * a.ncplx + (size_t)z * K1 * H0;
    for (int i = tid; i < K1 * H0; i += bd) buf[i] = in[i];
    for (int i = tid; i < K0 + K1; i += bd) w0[i] = a.tw[i];
    __syncthreads();
    cpx<T> acc[NT][FFT_R];
    dft_slow<T, NT, true>(buf, w1, K1, H0, tid, bd, acc);
    __syncthreads();
    dft_slow_store<T, NT>(buf, K1, H0, tid, bd, acc);
    __syncthreads();
    // x: half-complex to real; task = point x, rows y0 .. y0+R-1:
    // v(x) = sum_m wt(m) Re(c_m e^{+2 pi i m x / K0}), wt = 1 at m = 0 and
    // m = K0/2, else 2
    T out[NT][FFT_R];
    const int ngy = (K1 + FFT_R - 1) / FFT_R;
#pragma unroll
    for (int q = 0; q < NT; q++) {
#pragma unroll
        for (int r = 0; r < FFT_R; r++) out[q][r] = T(0);
        const int t = tid + q * bd;
        if (t >= K0 * ngy) continue;
        const int x = t % K0, y0 = (t / K0) * FFT_R;
        const cpx<T> *row[FFT_R];
#pragma unroll
        for (int r = 0; r < FFT_R; r++) row[r] = buf + (size_t)min(y0 + r, K1 - 1) * H0;
        int idx = 0;
        for (int m = 0; m < H0; m++) {
            const cpx<T> w = w0[idx];
            idx += x;
            if (idx >= K0) idx -= K0;
            const T wt = (m == 0 || 2 * m == K0) ? T(1) : T(2);
            const T wr = wt * w.re, wi = wt * w.im;
#pragma unroll
            for (int r = 0; r < FFT_R; r++) {
                const cpx<T> c = row[r][m];
                out[q][r] += c.re * wr + c.im * wi;
            }
        }
    }
    __syncthreads();
    T *rb = reinterpret_cast<T *>(buf);
#pragma unroll
    for (int q = 0; q < NT; q++) {
        const int t = tid + q * bd;
        if (t >= K0 * ngy) continue;
        const int x = t % K0, y0 = (t / K0) * FFT_R;
#pragma unroll
        for (int r = 0; r < FFT_R; r++)
            if (y0 + r < K1) rb[(size_t)(y0 + r) * K0 + x] = out[q][r];
    }
    __syncthreads();
    T *dst = a.real + set * a.nreal + (size_t)z * K0 * K1;
    for (int i = tid; i < K0 * K1; i += bd) dst[i] = rb[i];
}

// ---------------------------------------------------------------------------
// host side

static size_t plane_lds(const pmq_ctx *c, size_t w)
{
    const size_t H0 = c->K[0] / 2 + 1;
    return (H0 * c->K[1] + c->K[0] + c->K[1]) * 2 * w;
}

static size_t column_lds(const pmq_ctx *c, size_t w)
{
    const size_t H0 = c->K[0] / 2 + 1;
    return (H0 * c->K[2] + c->K[2]) * 2 * w;
}

bool fft_own_fits(const pmq_ctx *c, bool fp32)
{
    const size_t w = fp32 ? sizeof(float) : sizeof(double);
    const int H0 = c->K[0] / 2 + 1;
    const int cr0 = (H0 + FFT_R - 1) / FFT_R, cr1 = (c->K[1] + FFT_R - 1) / FFT_R;
    const int tasks = std::max({H0 * cr1, c->K[1] * cr0, c->K[2] * cr0, c->K[0] * cr1});
    return plane_lds(c, w) <= FFT_MAX_LDS && column_lds(c, w) <= FFT_MAX_LDS && tasks <= 4 * 1024;
}

// twiddle tables of the current grid, FP64 and FP32
int fft_own_setup(pmq_ctx *c)
{
    const int tot = c->K[0] + c->K[1] + c->K[2];
    std::vector<double> d(2 * (size_t)tot);
    std::vector<float> f(2 * (size_t)tot);
    int o = 0;
    for (int ax = 0; ax < 3; ax++) {
        const int K = c->K[ax];
        for (int j = 0; j < K; j++) {
            const double th = 2.0 * M_PI * (double)j / (double)K;
            d[2 * (o + j)] = cos(th);
            d[2 * (o + j) + 1] = -sin(th);
        }
        o += K;
    }
    for (size_t i = 0; i < d.size(); i++) f[i] = (float)d[i];
    PMQ_HIP(c->fft_tw.alloc(d.size()));
    PMQ_HIP(c->fft_tw_s.alloc(f.size()));
    PMQ_HIP(hipMemcpyAsync(c->fft_tw.p, d.data(), d.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PMQ_HIP(hipMemcpyAsync(c->fft_tw_s.p, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    PMQ_HIP(hipStreamSynchronize(c->stream)); // d, f are host temporaries
    return PMQ_OK;
}

// lanes per workgroup and outputs-groups per lane for `tasks` tasks
static void fft_shape(int tasks, int &nt, int &bd)
{
    nt = tasks <= 1024 ? 1 : tasks <= 2048 ? 2 : 4;
    bd = ((tasks + nt - 1) / nt + 63) / 64 * 64;
}

template <class T, int NT>
static void launch_fwd(const FftArgs<T> &a, int nset, int bd, size_t lds, hipStream_t st)
{
    hipLaunchKernelGGL((k_fft_planes_fwd<T, NT>), dim3(a.K2, nset), dim3(bd), lds, st, a);
}
template <class T, int NT>
static void launch_col(const FftArgs<T> &a, int nset, int bd, size_t lds, hipStream_t st)
{
    hipLaunchKernelGGL((k_fft_columns<T, NT>), dim3(a.K1, nset), dim3(bd), lds, st, a);
}
template <class T, int NT>
static void launch_inv(const FftArgs<T> &a, int nset, int bd, size_t lds, hipStream_t st)
{
    hipLaunchKernelGGL((k_fft_planes_inv<T, NT>), dim3(a.K2, nset), dim3(bd), lds, st, a);
}

#define PMQ_FFT_NT(FN, NT, ...)          \
    switch (NT) {                        \
    case 1: FN<T, 1>(__VA_ARGS__); break; \
    case 2: FN<T, 2>(__VA_ARGS__); break; \
    default: FN<T, 4>(__VA_ARGS__); break; \
    }

template <class T>
static int own_convolve(pmq_ctx *c, int nset, const T *G, const T *tw)
{
    FftArgs<T> a;
    a.K0 = c->K[0]; a.K1 = c->K[1]; a.K2 = c->K[2]; a.H0 = c->K[0] / 2 + 1;
    a.tw = reinterpret_cast<const cpx<T> *>(tw);
    a.real = reinterpret_cast<T *>(c->qgrid.p);
    a.cplx = reinterpret_cast<cpx<T> *>(c->cgrid.p);
    a.G = G;
    a.nreal = (size_t)a.K0 * a.K1 * a.K2;
    a.ncplx = (size_t)a.H0 * a.K1 * a.K2;
    const int ceilR0 = (a.H0 + FFT_R - 1) / FFT_R, ceilR1 = (a.K1 + FFT_R - 1) / FFT_R;
    const size_t w = sizeof(T);
    int nt, bd;
    // forward planes: x pass H0 x ceil(K1/R) tasks, y pass K1 x ceil(H0/R)
    fft_shape(std::max(a.H0 * ceilR1, a.K1 * ceilR0), nt, bd);
    PMQ_FFT_NT(launch_fwd, nt, a, nset, bd, plane_lds(c, w), c->stream)
    fft_shape(a.K2 * ceilR0, nt, bd);
    PMQ_FFT_NT(launch_col, nt, a, nset, bd, column_lds(c, w), c->stream)
    // inverse planes: y pass K1 x ceil(H0/R), x pass K0 x ceil(K1/R)
    fft_shape(std::max(a.K1 * ceilR0, a.K0 * ceilR1), nt, bd);
    PMQ_FFT_NT(launch_inv, nt, a, nset, bd, plane_lds(c, w), c->stream)
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

// qgrid (nset real grids) <- IFFT(G FFT(qgrid)), cgrid as the complex scratch
int fft_own_convolve(pmq_ctx *c, int nset, const void *G, bool fp32)
{
    if (!fft_own_fits(c, fp32)) {
        c->err = "PME: the grid planes do not fit the LDS of the direct transforms";
        return PMQ_ERR_UNSUPPORTED;
    }
    if (fp32) return own_convolve<float>(c, nset, (const float *)G, c->fft_tw_s.p);
    return own_convolve<double>(c, nset, (const double *)G, c->fft_tw.p);
}

} // namespace pmq
