// k_fft.hip -- the PME convolution (forward 3-D real FFT, multiplication by
// the reciprocal kernel G, inverse FFT) for grids whose planes fit in LDS,
// in three launches instead of six library transforms and a separate
// multiplication:
//   k_fft_planes_fwd  per z-plane, grid set and group of MW half-complex
//                     columns m: the real-to-complex DFT along x of those
//                     columns, then their complex DFT along y
//   k_fft_columns     per y-row, grid set and group of MW columns m: DFT
//                     along z, multiplication by G, inverse DFT along z
//   k_fft_planes_inv  per z-plane, grid set and group of YW rows y: the
//                     inverse DFT along y of those rows (every column),
//                     then complex-to-real along x
// Every output is computed by exactly one workgroup, so splitting the planes
// into column or row groups costs no arithmetic; it only has each workgroup
// read the whole plane, which it does from L2.  A 48^3 grid with two sets
// runs about 500 workgroups per launch rather than 96 whole planes.
// Every transform is a direct DFT sum (each output a length-K dot product
// with a twiddle table held in LDS).  A group of S lanes computes R outputs
// of one frequency, each lane a 1/S share of the terms, so the inputs it
// reads are shared with the groups of the other frequencies (LDS
// broadcasts) and only the twiddle differs between groups.
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
constexpr int FFT_MW = 5;           // half-complex columns per workgroup (forward planes, z columns)
constexpr int FFT_YW = 8;           // rows per workgroup (inverse planes)
constexpr int FFT_MAX_LDS = 65536;  // bytes of LDS a workgroup may take
constexpr int FFT_MAX_TB = 1024;    // lanes per workgroup

template <class T>
struct FftArgs {
    int K0, K1, K2, H0;
    const cpx<T> *tw; // forward twiddles e^{-2 pi i j/K} of axes 0, 1, 2 (K0 + K1 + K2 entries)
    T *real;          // nset real grids, nreal apart
    cpx<T> *cplx;     // nset complex half grids, ncplx apart
    const T *G;       // H0 K1 K2 reciprocal kernel
    size_t nreal, ncplx;
    const double *skip; // nonzero: return at once (or null)
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

// Each sum is split over FFT_S consecutive lanes: lane `sub` of the group
// adds the terms j = sub, sub + S, ... and the group adds its partials with
// butterfly shuffles (a fixed order), which cuts the chain of dependent LDS
// reads per lane by S.  All lanes of a group take the same loop trips.
constexpr int FFT_S = 4;

template <class T>
__device__ __forceinline__ T group_sum(T v)
{
#pragma unroll
    for (int o = FFT_S / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Complex DFT along the slow index of an L x W array in LDS (row j holds W
// values, buf[j*W + m]) for output frequency k and columns m0 .. m0+R-1
// (clamped to W-1; the caller drops those): acc[r] = sum_j buf[j*W + m]
// tw[(k j) mod L], conjugated twiddles for the inverse.
template <class T, bool INV>
__device__ __forceinline__ void dft_column(const cpx<T> *buf, const cpx<T> *tw, int L, int W, int k, int m0, int sub,
                                           cpx<T> (&acc)[FFT_R])
{
    int col[FFT_R];
#pragma unroll
    for (int r = 0; r < FFT_R; r++) {
        acc[r] = cpx<T>{T(0), T(0)};
        col[r] = min(m0 + r, W - 1);
    }
    int idx = (k * sub) % L;
    const int step = (k * FFT_S) % L;
    for (int j = sub; j < L; j += FFT_S) {
        const cpx<T> w = tw[idx];
        idx += step;
        if (idx >= L) idx -= L;
        const cpx<T> *row = buf + (size_t)j * W;
#pragma unroll
        for (int r = 0; r < FFT_R; r++) {
            if (INV) cmadc(acc[r], row[col[r]], w);
            else cmad(acc[r], row[col[r]], w);
        }
    }
#pragma unroll
    for (int r = 0; r < FFT_R; r++) {
        acc[r].re = group_sum(acc[r].re);
        acc[r].im = group_sum(acc[r].im);
    }
}

// global -> LDS copy of n elements by the workgroup, U loads per lane in
// flight before their stores
template <class V, int U = 8>
__device__ __forceinline__ void copy_in(V *dst, const V *src, int n, int tid, int bd)
{
    for (int base = 0; base < n; base += bd * U) {
        V tmp[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = base + u * bd + tid;
            if (i < n) tmp[u] = src[i];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = base + u * bd + tid;
            if (i < n) dst[i] = tmp[u];
        }
    }
}

template <class T>
__global__ __launch_bounds__(FFT_MAX_TB) void k_fft_planes_fwd(FftArgs<T> a)
{
    if (a.skip && *a.skip != 0.0) return;
    extern __shared__ double fft_lds[];
    const int K0 = a.K0, K1 = a.K1, H0 = a.H0;
    const int z = blockIdx.x, set = blockIdx.y, tid = threadIdx.x, bd = blockDim.x;
    const int sub = tid % FFT_S, gt = tid / FFT_S, ng = bd / FFT_S;
    const int m0 = blockIdx.z * FFT_MW, mw = min(FFT_MW, H0 - m0);
    T *rb = reinterpret_cast<T *>(fft_lds);                                      // real plane K1 x K0
    cpx<T> *cb = reinterpret_cast<cpx<T> *>(fft_lds + ((size_t)K0 * K1 * sizeof(T) + 15) / 16 * 2); // K1 x mw
    cpx<T> *w0 = cb + (size_t)K1 * FFT_MW, *w1 = w0 + K0;
    const T *in = a.real + set * a.nreal + (size_t)z * K0 * K1;
    copy_in(rb, in, K0 * K1, tid, bd);
    copy_in(w0, a.tw, K0 + K1, tid, bd);
    __syncthreads();
    // x: real to the group's half-complex columns; task = column mi, rows y0 .. y0+R-1
    const int ngy = (K1 + FFT_R - 1) / FFT_R;
    for (int t = gt; t < mw * ngy; t += ng) {
        const int mi = t % mw, y0 = (t / mw) * FFT_R, m = m0 + mi;
        const T *row[FFT_R];
#pragma unroll
        for (int r = 0; r < FFT_R; r++) row[r] = rb + (size_t)min(y0 + r, K1 - 1) * K0;
        cpx<T> acc[FFT_R];
#pragma unroll
        for (int r = 0; r < FFT_R; r++) acc[r] = cpx<T>{T(0), T(0)};
        int idx = (m * sub) % K0;
        const int step = (m * FFT_S) % K0;
        for (int x = sub; x < K0; x += FFT_S) {
            const cpx<T> w = w0[idx];
            idx += step;
            if (idx >= K0) idx -= K0;
#pragma unroll
            for (int r = 0; r < FFT_R; r++) {
                const T d = row[r][x];
                acc[r].re += d * w.re;
                acc[r].im += d * w.im;
            }
        }
#pragma unroll
        for (int r = 0; r < FFT_R; r++) {
            acc[r].re = group_sum(acc[r].re);
            acc[r].im = group_sum(acc[r].im);
            if (sub == r && y0 + r < K1) cb[(size_t)(y0 + r) * mw + mi] = acc[r];
        }
    }
    __syncthreads();
    // y: complex DFT of the group's columns, written to the half grid
    cpx<T> *out = a.cplx + set * a.ncplx + (size_t)z * K1 * H0;
    const int ngm = (mw + FFT_R - 1) / FFT_R;
    for (int t = gt; t < K1 * ngm; t += ng) {
        const int k = t % K1, mi0 = (t / K1) * FFT_R;
        cpx<T> acc[FFT_R];
        dft_column<T, false>(cb, w1, K1, mw, k, mi0, sub, acc);
#pragma unroll
        for (int r = 0; r < FFT_R; r++)
            if (sub == r && mi0 + r < mw) out[(size_t)k * H0 + m0 + mi0 + r] = acc[r];
    }
}

template <class T>
__global__ __launch_bounds__(FFT_MAX_TB) void k_fft_columns(FftArgs<T> a)
{
    if (a.skip && *a.skip != 0.0) return;
    extern __shared__ double fft_lds[];
    const int K1 = a.K1, K2 = a.K2, H0 = a.H0;
    const int y = blockIdx.x, set = blockIdx.y, tid = threadIdx.x, bd = blockDim.x;
    const int sub = tid % FFT_S, gt = tid / FFT_S;
    const int m0 = blockIdx.z * FFT_MW, mw = min(FFT_MW, H0 - m0);
    cpx<T> *buf = reinterpret_cast<cpx<T> *>(fft_lds); // K2 x mw: row z
    cpx<T> *w2 = buf + (size_t)K2 * FFT_MW;
    cpx<T> *cg = a.cplx + set * a.ncplx;
    {
        constexpr int U = 8;
        for (int base = 0; base < K2 * mw; base += bd * U) {
            cpx<T> tmp[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = base + u * bd + tid;
                if (i < K2 * mw) {
                    const int z = i / mw, mi = i - z * mw;
                    tmp[u] = cg[((size_t)z * K1 + y) * H0 + m0 + mi];
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = base + u * bd + tid;
                if (i < K2 * mw) buf[i] = tmp[u];
            }
        }
    }
    copy_in(w2, a.tw + a.K0 + K1, K2, tid, bd);
    __syncthreads();
    const int ngm = (mw + FFT_R - 1) / FFT_R;
    const int ntask = K2 * ngm;
    // forward along z, times G(m, y, k3); results wait in registers for the sync
    cpx<T> acc[FFT_R];
    const bool act = gt < ntask; // the same for the lanes of a group
    const int k = act ? gt % K2 : 0, mi0 = act ? (gt / K2) * FFT_R : 0;
    if (act) dft_column<T, false>(buf, w2, K2, mw, k, mi0, sub, acc);
    __syncthreads();
    if (act) {
#pragma unroll
        for (int r = 0; r < FFT_R; r++) {
            const int mi = mi0 + r;
            if (sub != r || mi >= mw) continue;
            const T g = a.G[((size_t)k * K1 + y) * H0 + m0 + mi];
            buf[(size_t)k * mw + mi] = cpx<T>{acc[r].re * g, acc[r].im * g};
        }
    }
    __syncthreads();
    // inverse along z, written back in place
    if (act) {
        dft_column<T, true>(buf, w2, K2, mw, k, mi0, sub, acc);
#pragma unroll
        for (int r = 0; r < FFT_R; r++)
            if (sub == r && mi0 + r < mw) cg[((size_t)k * K1 + y) * H0 + m0 + mi0 + r] = acc[r];
    }
}

template <class T>
__global__ __launch_bounds__(FFT_MAX_TB) void k_fft_planes_inv(FftArgs<T> a)
{
    if (a.skip && *a.skip != 0.0) return;
    extern __shared__ double fft_lds[];
    const int K0 = a.K0, K1 = a.K1, H0 = a.H0;
    const int z = blockIdx.x, set = blockIdx.y, tid = threadIdx.x, bd = blockDim.x;
    const int sub = tid % FFT_S, gt = tid / FFT_S, ng = bd / FFT_S;
    const int y0 = blockIdx.z * FFT_YW, yw = min(FFT_YW, K1 - y0);
    cpx<T> *cb = reinterpret_cast<cpx<T> *>(fft_lds); // the complex plane, K1 x H0
    cpx<T> *ob = cb + (size_t)K1 * H0;                 // the group's rows after the y transform, yw x H0
    cpx<T> *w0 = ob + (size_t)FFT_YW * H0, *w1 = w0 + K0;
    const cpx<T> *in = a.cplx + set * a.ncplx + (size_t)z * K1 * H0;
    copy_in(cb, in, K1 * H0, tid, bd);
    copy_in(w0, a.tw, K0 + K1, tid, bd);
    __syncthreads();
    // y: inverse DFT onto the group's rows; task = row yi, columns m0 .. m0+R-1
    const int ngm = (H0 + FFT_R - 1) / FFT_R;
    for (int t = gt; t < yw * ngm; t += ng) {
        const int yi = t % yw, m0 = (t / yw) * FFT_R;
        cpx<T> acc[FFT_R];
        dft_column<T, true>(cb, w1, K1, H0, y0 + yi, m0, sub, acc);
#pragma unroll
        for (int r = 0; r < FFT_R; r++)
            if (sub == r && m0 + r < H0) ob[(size_t)yi * H0 + m0 + r] = acc[r];
    }
    __syncthreads();
    // x: half-complex to real; task = point x, rows yi0 .. yi0+R-1:
    // v(x) = sum_m wt(m) Re(c_m e^{+2 pi i m x / K0}), wt = 1 at m = 0 and
    // m = K0/2, else 2
    T *dst = a.real + set * a.nreal + (size_t)z * K0 * K1;
    const int ngy = (yw + FFT_R - 1) / FFT_R;
    for (int t = gt; t < K0 * ngy; t += ng) {
        const int x = t % K0, yi0 = (t / K0) * FFT_R;
        const cpx<T> *row[FFT_R];
#pragma unroll
        for (int r = 0; r < FFT_R; r++) row[r] = ob + (size_t)min(yi0 + r, yw - 1) * H0;
        T out[FFT_R];
#pragma unroll
        for (int r = 0; r < FFT_R; r++) out[r] = T(0);
        int idx = (x * sub) % K0;
        const int step = (x * FFT_S) % K0;
        for (int m = sub; m < H0; m += FFT_S) {
            const cpx<T> w = w0[idx];
            idx += step;
            if (idx >= K0) idx -= K0;
            const T wt = (m == 0 || 2 * m == K0) ? T(1) : T(2);
            const T wr = wt * w.re, wi = wt * w.im;
#pragma unroll
            for (int r = 0; r < FFT_R; r++) {
                const cpx<T> c = row[r][m];
                out[r] += c.re * wr + c.im * wi;
            }
        }
#pragma unroll
        for (int r = 0; r < FFT_R; r++) {
            out[r] = group_sum(out[r]);
            if (sub == r && yi0 + r < yw) dst[(size_t)(y0 + yi0 + r) * K0 + x] = out[r];
        }
    }
}

// ---------------------------------------------------------------------------
// host side

static size_t fwd_lds(const pmq_ctx *c, size_t w)
{
    const size_t real = ((size_t)c->K[0] * c->K[1] * w + 15) / 16 * 16;
    return real + ((size_t)c->K[1] * FFT_MW + c->K[0] + c->K[1]) * 2 * w;
}

static size_t col_lds(const pmq_ctx *c, size_t w) { return ((size_t)c->K[2] * FFT_MW + c->K[2]) * 2 * w; }

static size_t inv_lds(const pmq_ctx *c, size_t w)
{
    const size_t H0 = c->K[0] / 2 + 1;
    return (H0 * c->K[1] + H0 * FFT_YW + c->K[0] + c->K[1]) * 2 * w;
}

static int round_tb(int tasks) { return std::max(64, (tasks * FFT_S + 63) / 64 * 64); }

// lanes of the three kernels (FFT_S lanes per task)
static void fft_tasks(const pmq_ctx *c, int tb[3])
{
    const int H0 = c->K[0] / 2 + 1;
    const int ngy = (c->K[1] + FFT_R - 1) / FFT_R, ngm = (FFT_MW + FFT_R - 1) / FFT_R;
    tb[0] = round_tb(std::max(FFT_MW * ngy, c->K[1] * ngm));
    tb[1] = round_tb(c->K[2] * ngm);
    tb[2] = round_tb(std::max(FFT_YW * ((H0 + FFT_R - 1) / FFT_R), c->K[0] * ((FFT_YW + FFT_R - 1) / FFT_R)));
}

bool fft_own_fits(const pmq_ctx *c, bool fp32)
{
    const size_t w = fp32 ? sizeof(float) : sizeof(double);
    int tb[3];
    fft_tasks(c, tb);
    return fwd_lds(c, w) <= FFT_MAX_LDS && col_lds(c, w) <= FFT_MAX_LDS && inv_lds(c, w) <= FFT_MAX_LDS &&
           std::max({tb[0], tb[1], tb[2]}) <= FFT_MAX_TB;
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

template <class T>
static int own_convolve(pmq_ctx *c, int nset, const T *G, const T *tw, const double *skip)
{
    FftArgs<T> a;
    a.K0 = c->K[0]; a.K1 = c->K[1]; a.K2 = c->K[2]; a.H0 = c->K[0] / 2 + 1;
    a.tw = reinterpret_cast<const cpx<T> *>(tw);
    a.real = reinterpret_cast<T *>(c->qgrid.p);
    a.cplx = reinterpret_cast<cpx<T> *>(c->cgrid.p);
    a.G = G;
    a.skip = skip;
    a.nreal = (size_t)a.K0 * a.K1 * a.K2;
    a.ncplx = (size_t)a.H0 * a.K1 * a.K2;
    const size_t w = sizeof(T);
    int tb[3];
    fft_tasks(c, tb);
    const unsigned mg = (unsigned)((a.H0 + FFT_MW - 1) / FFT_MW), yg = (unsigned)((a.K1 + FFT_YW - 1) / FFT_YW);
    hipLaunchKernelGGL(k_fft_planes_fwd<T>, dim3(a.K2, nset, mg), dim3(tb[0]), fwd_lds(c, w), c->stream, a);
    hipLaunchKernelGGL(k_fft_columns<T>, dim3(a.K1, nset, mg), dim3(tb[1]), col_lds(c, w), c->stream, a);
    hipLaunchKernelGGL(k_fft_planes_inv<T>, dim3(a.K2, nset, yg), dim3(tb[2]), inv_lds(c, w), c->stream, a);
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

// qgrid (nset real grids) <- IFFT(G FFT(qgrid)), cgrid as the complex scratch
int fft_own_convolve(pmq_ctx *c, int nset, const void *G, bool fp32, const double *skip)
{
    if (!fft_own_fits(c, fp32)) {
        c->err = "PME: the grid planes do not fit the LDS of the direct transforms";
        return PMQ_ERR_UNSUPPORTED;
    }
    if (fp32) return own_convolve<float>(c, nset, (const float *)G, c->fft_tw_s.p, skip);
    return own_convolve<double>(c, nset, (const double *)G, c->fft_tw.p, skip);
}

} // namespace pmq
