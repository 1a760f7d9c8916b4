// pmq_internal.h -- shared definitions of the HIP library behind include/pmq.h.
//
// Data layout in HBM (one context = one system on one GPU):
//   * coordinates x,y,z are kept in the caller's atom order (double[n] each);
//   * every eval sorts atoms spatially (cells in Hilbert order, stable by atom
//     index) and the pair kernels work on "sorted" SoA copies: positions,
//     the 10 rotated multipole components, and per-atom flags, padded to a
//     multiple of 64 (one wavefront per 64-atom block);
//   * pair work is a list of (I,J) block tiles, I <= J, whose bounding boxes
//     come within the cutoff; each tile is one wavefront;
//   * forces, torques and per-atom potentials are accumulated in the
//     caller's atom order, so the frame (torque) and charge-flux kernels
//     index frame atoms directly.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/pmq.h"

#define PMQ_BLOCK 64
#define PMQ_BLK_STRIDE 12 // doubles per block in the block-bounds array
#define PMQ_PK 10         // doubles per atom in a packed product input: x y z, two dipole sets, pad
#define PMQ_NOALLOC_MAX_DEVICES 64

namespace pmq {

// Slot of the kept block pair (I, J), I <= J, in the tile list: the nblk
// diagonal tiles (always kept: a block is within the cutoff of itself) come
// first, in block order, then the others in (I, J) order.  scan is the
// pair's exclusive count of kept pairs in (I, J) order, which before an
// off-diagonal pair of row I includes the I + 1 diagonals of rows 0..I.
// The tile kernels take tiles in slot order, one wave each, and a diagonal
// tile holds several times the pairs of the others (about 1,200 against 190
// on a water box with 64-atom Hilbert blocks and a 7 A cutoff), so issuing
// them first keeps them off the end of the launch.
__host__ __device__ __forceinline__ int tile_slot(int I, int J, int nblk, int scan)
{
    return I == J ? I : nblk + scan - (I + 1);
}

// ---------------------------------------------------------------------------
// periodic box as the kernels see it (lattice constants of lattice.cpp)
struct DevBox {
    int kind;
    double xbox, ybox, zbox;
    double beta_sin, beta_cos, gamma_sin, gamma_cos, beta_term, gamma_term;
    double recip[3][3]; // fractional = recip * cartesian (rows a*, b*, c*)
    double lvec[3][3];  // cartesian = lvec * fractional (columns a, b, c)
};

DevBox make_box(const pmq_box *b);

// minimum image of one displacement, the reference's image() (image.cpp:37-124)
template <class R>
__device__ __host__ inline void image_wrap(R &v, R cell)
{
    const R half = R(0.5) * cell;
    while (fabs(v) > half) v -= copysign(cell, v);
}

template <class R>
__device__ __host__ inline void image(const DevBox &b, R &x, R &y, R &z)
{
    switch (b.kind) {
    case PMQ_BOX_ORTHO:
        image_wrap<R>(x, R(b.xbox)); image_wrap<R>(y, R(b.ybox)); image_wrap<R>(z, R(b.zbox));
        break;
    case PMQ_BOX_MONO:
        z = z / R(b.beta_sin);
        x = x - z * R(b.beta_cos);
        image_wrap<R>(x, R(b.xbox)); image_wrap<R>(y, R(b.ybox)); image_wrap<R>(z, R(b.zbox));
        x = x + z * R(b.beta_cos);
        z = z * R(b.beta_sin);
        break;
    case PMQ_BOX_TRI:
        z = z / R(b.gamma_term);
        y = (y - z * R(b.beta_term)) / R(b.gamma_sin);
        x = x - y * R(b.gamma_cos) - z * R(b.beta_cos);
        image_wrap<R>(x, R(b.xbox)); image_wrap<R>(y, R(b.ybox)); image_wrap<R>(z, R(b.zbox));
        x = x + y * R(b.gamma_cos) + z * R(b.beta_cos);
        y = y * R(b.gamma_sin) + z * R(b.beta_term);
        z = z * R(b.gamma_term);
        break;
    case PMQ_BOX_OCTA:
        image_wrap<R>(x, R(b.xbox)); image_wrap<R>(y, R(b.ybox)); image_wrap<R>(z, R(b.zbox));
        if (fabs(x) + fabs(y) + fabs(z) > R(0.75) * R(b.xbox)) {
            x -= copysign(R(0.5) * R(b.xbox), x);
            y -= copysign(R(0.5) * R(b.ybox), y);
            z -= copysign(R(0.5) * R(b.zbox), z);
        }
        break;
    case PMQ_BOX_DODECA: {
        const R root2 = R(1.41421356237309504880);
        image_wrap<R>(x, R(b.xbox)); image_wrap<R>(y, R(b.ybox));
        z = z - root2 * R(b.zbox) * round(z / (R(b.zbox) * root2));
        const R corr = R(0.5) * R(b.xbox) * R((int)(fabs(x / R(b.xbox)) + fabs(y / R(b.ybox)) + fabs(root2 * z / R(b.zbox))));
        x -= copysign(corr, x);
        y -= copysign(corr, y);
        z -= copysign(corr, z) * root2;
        break;
    }
    default:
        break;
    }
}

// ---------------------------------------------------------------------------
// device buffers
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    hipError_t alloc(size_t count)
    {
        if (count <= n && p) return hipSuccess;
        release();
        const hipError_t e = hipMalloc((void **)&p, (count ? count : 1) * sizeof(T));
        if (e == hipSuccess) n = count;
        else p = nullptr;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// mixed precision accumulates gradients, torques and fields in 64-bit fixed
// point: integer atomics are order independent, so the sums (and everything
// downstream) are bit-reproducible run to run.  Scales: 2^32 per kcal/mol/A
// (range 2^31), 2^40 per field unit (range 2^23).
constexpr double PMQ_FX_GRAD = 4294967296.0;
constexpr double PMQ_FX_FIELD = 1099511627776.0;
constexpr double PMQ_FX_CF = 1125899906842624.0; // 2^50 per charge: the charge-flux monopole shifts

__device__ __forceinline__ void fx_add(long long *q, size_t i, double v, double scale)
{
    atomicAdd((unsigned long long *)(q + i), (unsigned long long)__double2ll_rn(v * scale));
}

// gradient / torque accumulation: FP64 atomics (plain form) or fixed point
__device__ __forceinline__ void grad_add(double *d, long long *q, size_t i, double v)
{
    if (q) fx_add(q, i, v, PMQ_FX_GRAD);
    else atomicAdd(d + i, v);
}

// what the host must check after an evaluation enqueued without round trips:
// the true tile and pair-matrix sizes and whether either exceeded the
// capacity the evaluation ran with
struct RunStatus {
    int ntiles, nnz, overflow, pad;
};
struct HostStatus {       // pinned host mirror, written by the evaluation's last copies
    RunStatus st;
    double sc[11];        // PCG scalars (8 = converged, 9 = iterations)
};

// Symmetrized pair virial, the reference's convention (empole.cpp:310-320,
// pairMpole.h:166-173): for a pair gradient g on the first atom at
// separation r = x_k - x_i, W_ab -= (r_a g_b + r_b g_a) / 2.  Six FP64
// accumulators whatever the pair arithmetic is.
struct Vir6 {
    double xx = 0, xy = 0, xz = 0, yy = 0, yz = 0, zz = 0;
    __device__ __forceinline__ void pair(double x, double y, double z, double gx, double gy, double gz)
    {
        xx -= x * gx;
        xy -= 0.5 * (y * gx + x * gy);
        xz -= 0.5 * (z * gx + x * gz);
        yy -= y * gy;
        yz -= 0.5 * (z * gy + y * gz);
        zz -= z * gz;
    }
    // sums of the 64 lanes of a wavefront, in every lane
    __device__ __forceinline__ void wave_sum()
    {
        double *v[6] = {&xx, &xy, &xz, &yy, &yz, &zz};
#pragma unroll
        for (int q = 0; q < 6; q++)
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) *v[q] += __shfl_xor(*v[q], o, 64);
    }
    // tile-item scalars 1..9 (row-major 3x3) of item t, scalar q at tscal[q * ts + t]
    __device__ __forceinline__ void store_tile(double *tscal, size_t ts, int t) const
    {
        double *S = tscal + t;
        S[ts] = xx; S[2 * ts] = xy; S[3 * ts] = xz; S[4 * ts] = xy; S[5 * ts] = yy; S[6 * ts] = yz;
        S[7 * ts] = xz; S[8 * ts] = yz; S[9 * ts] = zz;
    }
};

// per-item partial sums of the pair kernels (k_pairs.hip): values per lane
// and side, scalars per item (0 energy, 1..9 virial, 10 inter-molecular
// energy, 11 pair count), stored scalar-major (scalar q of item u at
// q * item capacity + u) so each scalar's sum reads one contiguous array.
// An item is one piece of a tile (tile_pairs.h TileItem): a tile with many
// pairs is split into up to PMQ_TILE_PIECES pieces of about
// PMQ_PIECE_BATCHES batches of 64 pairs, each run by its own wave.
constexpr int PMQ_TILE_NV = 8;
constexpr int PMQ_TILE_NS = 12;
constexpr int PMQ_TILE_PIECES = 4;
constexpr int PMQ_PIECE_BATCHES = 6;
constexpr int PMQ_SPLIT_TILES = 16384; // tile lists longer than this are not split
constexpr int PMQ_SHP_ROWS = 8; // rows of the sorted per-atom HIPPO, charge-transfer and activity table

struct TileReduceArgs {
    int nblk, n;
    const int2 *tiles;
    const int *adj, *adjn, *perm;
    const int *tpc, *ntiles; // piece code of each tile (k_tile_masks), tile count
    const void *part;
    double *out[8];       // per value: destination (or null)
    double *out2[8];      // second destination of the same value (or null)
    int caller_order[8];  // index by caller order (perm) or by sorted order
    int stride[8], off[8];
    double fxscale[8];    // fixed-point scale of each value (mixed form)
    // fixed-point accumulators (same index as out / out2) whose values join
    // the sums and are cleared: the scaled-pair kernels' shares (mixed
    // form; null: none)
    long long *fxq[8], *fxq2[8];
    // with k1 > k0 one more workgroup sums the tile-item scalars [k0, k1)
    // into the totals (tile_scalars), beside the per-atom sums
    int scal_which, scal_k0, scal_k1;
    const double *tscal;
    size_t tstride;
    struct Totals *tot;
};

// energy / virial / count accumulators (one per context, in device memory)
struct Totals {
    double em, ep, er, edsp, ev, ect; // indexed 0..5 from em by tile_scalars
    double einter;
    unsigned long long nem;
    double vir[9];
    __device__ __forceinline__ void add_vir(const Vir6 &v)
    {
        atomicAdd(&vir[0], v.xx); atomicAdd(&vir[4], v.yy); atomicAdd(&vir[8], v.zz);
        atomicAdd(&vir[1], v.xy); atomicAdd(&vir[3], v.xy);
        atomicAdd(&vir[2], v.xz); atomicAdd(&vir[6], v.xz);
        atomicAdd(&vir[5], v.yz); atomicAdd(&vir[7], v.yz);
    }
    // fixed-point accumulators of the mixed path
    long long em_fx, ep_fx;
    long long vir_fx[9];
    int pcg_iters;
    int pad;
    unsigned long long nep; // analysis: polarizable sites with a nonzero energy
};

// a pair whose interactions are scaled (connectivity 1-2..1-5 and
// polarization groups 1-1..1-4): the class codes are fixed by the topology,
// the scale values come with the parameters
struct ExclPair {
    int i, k;       // original atom indices, i < k
    int mcls;       // 2..5 bonded class for m/p scaling, 0 = none
    int gcls;       // 1..4 polarization-group class for d/u scaling, 0 = none
    int intra;      // 1 if k is in i's own polarization group (pNiscale)
    int in14;       // 1 if k is listed 1-4 of i (the vdW 1-4 tables), whatever mcls says
    double m, p, d, u; // scale values (filled by set_params); u is the mutual-field scale (w for HIPPO)
    double dsp, rep;   // HIPPO dispersion and repulsion scales (by connectivity)
    double v;          // van der Waals scale (by connectivity)
};

} // namespace pmq

// ---------------------------------------------------------------------------
struct pmq_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // sizes
    int n = 0;      // atoms
    int npad = 0;   // atoms padded to a whole number of 64-atom blocks
    int nblk = 0;   // blocks
    bool have_top = false, have_prm = false, have_xyz = false;

    // ---- topology (host copies kept for the exclusion list)
    std::vector<int> h_polaxe, h_zaxis, h_xaxis, h_yaxis, h_molcule, h_usef;
    std::vector<int> h_use;          // n: use[] itself (without the frame atoms usef adds)
    std::vector<pmq::ExclPair> h_excl;
    int nbond = 0, nangle = 0;

    // ---- device: static per topology / params
    pmq::DBuf<int> polaxe, zaxis, xaxis, yaxis, molcule, usef, use_raw;
    pmq::DBuf<pmq::ExclPair> excl;
    pmq::DBuf<double> pole;          // 13n local
    pmq::DBuf<double> pcore, pval0, palpha;
    pmq::DBuf<int> ibnd, iang, balist;
    pmq::DBuf<double> bl, bflx, anat, aflx, abflx;
    pmq_params prm{};                // scalars only (pointers cleared)

    // ---- device: per eval
    pmq::DBuf<double> x, y, z;
    pmq::DevBox box{};
    pmq::DBuf<double> pdelta;        // charge-flux monopole shift, as 64-bit fixed point (PMQ_FX_CF)
    pmq::DBuf<double> potp;          // charge flux: dE_pol/dq per atom (caller order)
    pmq::DBuf<long long> fx_pot;     // mixed form: 2n fixed-point dE/dq (multipoles, polarization) of the scaled pairs
    pmq::DBuf<double> rpole13;       // 13n global-frame multipoles, caller order
    // sorted SoA
    pmq::DBuf<int> perm, iperm;      // sorted -> original, original -> sorted
    pmq::DBuf<double> sx, sy, sz;    // npad each
    pmq::DBuf<double> smp;           // 10 * npad rotated multipoles (charge = valence if CP)
    pmq::DBuf<double> scp;           // 3 * npad: core, alpha (CP), spare
    pmq::DBuf<int> sflag;            // npad: bit0 valid, bit1 active(use)
    pmq::DBuf<int> smol;             // npad molecule ids
    // neighbor structure
    pmq::DBuf<unsigned long long> keys, keys_sorted;
    pmq::DBuf<double> blk;           // PMQ_BLK_STRIDE * nblk: center xyz, half extent xyz (wrapped into the
                                     // cell), then lowest and highest stored coordinate xyz
    pmq::DBuf<int> tileflag, tilescan;
    pmq::DBuf<int2> tiles;
    pmq::DBuf<int> ntiles_d;                     // [0] tiles listed, [1] further pieces (xlist entries)
    pmq::DBuf<double> bbox;          // 6: min xyz, max xyz (non-periodic systems)
    pmq::DBuf<unsigned char> scratch_sort;
    size_t scratch_sort_bytes = 0;
    int ntiles = 0;
    int ntiles_cap = 0;
    // outputs
    pmq::DBuf<double> dem, te, pot, aem, aep;
    pmq::DBuf<long long> fx_grad;    // 6n fixed-point gradient + torque (mixed form), zero between uses
    pmq::DBuf<double> rotm;          // 9n local-frame rotation of each site (rows x, y, z)
    pmq::DBuf<int> flip;             // n chirality flips of the last frame pass
    // parameter gradients (k_params.hip): sorted dE/d(global multipole), dE/d(pdamp), dE/d(thole)
    pmq::DBuf<double> pg_dm, pg_dpd, pg_dth;
    pmq::DBuf<long long> fx_field;   // 6 npad fixed-point fields (mixed form), zero between uses
    pmq::DBuf<pmq::Totals> totals;

    // ---- polarization and PME state (pmq_polar.cpp)
    pmq::DBuf<double> polarity, thole, pdamp;    // caller order
    std::vector<double> h_polarity;
    pmq::DBuf<double> spol;                      // 3 * npad sorted: polarity, thole, pdamp
    pmq::DBuf<double> cphi;                      // 10 * npad sorted: phi, grad(3), hessian(6) of the
                                                 // permanent multipoles' reciprocal potential
    // two-set vectors (set 0 = d, set 1 = p), each 2 x 3 x npad sorted SoA
    pmq::DBuf<double> pfield;                    // permanent fields E^d, E^p
    pmq::DBuf<double> uind;                      // induced dipoles mu_d, mu_p
    pmq::DBuf<double> pres, pz, pdir, pvec;      // PCG residual, preconditioned residual,
                                                 // search direction, T applied to the operator's input
    pmq::DBuf<double> pav;                       // A p = (1/alpha - T) p
    pmq::DBuf<double> pcg;                       // PCG scalars
    pmq::DBuf<double> pcg_part;                  // PCG per-block partial sums
    size_t pcg_part1_cap = 0;                    // E1 partials that fit before the E2 ones in pcg_part
    int pcg_guess_iters = 0;                     // iterations of the last converged solve (0 = none yet)
    pmq::DBuf<double> tep;                       // 3n torques of the polarization term
    pmq::DBuf<double> dep;                       // 3n gradient of ep
    bool have_uind = false;
    // dipole-field operator T as a row-owned sparse matrix (k_pairs.hip)
    pmq::DBuf<unsigned long long> tmask;         // 128 per tile: 64 row masks, 64 column masks
    pmq::DBuf<unsigned long long> xmask;         // 128 per tile: tmask rows and mmask rows without the scaled pairs
    pmq::DBuf<unsigned long long> mmask;         // 64 per tile: row masks of pairs with an active atom
    pmq::DBuf<double> tpart;                     // per tile item: 2 sides x PMQ_TILE_NV x 64 partial sums
    pmq::DBuf<double> tscal;                     // per tile item: PMQ_TILE_NS scalars (energy, virial, ...)
    pmq::DBuf<int> tpc;                          // per tile: piece count | batches per piece << 4
    pmq::DBuf<int> xlist;                        // further pieces of the split tiles (tile | piece << 28)
    int tgrid = 1;                               // workgroups of the per-tile kernels (masks, parameters)
    int igrid = 1;                               // workgroups of the per-item pair kernels
    int tile_cap = 0;                            // tile capacity of the tile buffers (0 = not sized yet)
    pmq::DBuf<int> ntiles_eff;                   // min(tile count, tile_cap): what the tile kernels loop over
    pmq::DBuf<pmq::RunStatus> status;
    pmq::HostStatus *h_status = nullptr;         // pinned
    pmq::HostStatus *h_status_dev = nullptr;     // its device address
    bool safe = true;                            // next evaluation sizes buffers and the PCG with host round trips
    bool pending = false;                        // an evaluation awaits its check (pmq_sync / pmq_fetch)
    bool pending_safe = false;                   // ... and it ran in safe mode
    pmq::DBuf<int> soff;                         // 128 per tile: each lane's entry offset within its row
    pmq::DBuf<int> adj, adjn;                    // per block: tiles it takes part in (nblk x nblk), count
    pmq::DBuf<int> rowcnt, rowptr;               // npad + 1
    pmq::DBuf<unsigned char> pscan_tmp;
    pmq::DBuf<double> tval;                      // tcap entries (R1, R2) of T = R2 r r^T - R1 I (double2, or float2 mixed)
    pmq::DBuf<int> tcol;                         // tcap columns of those entries (shared with tval_s)
    pmq::DBuf<double> pk_src, pk_w;              // PMQ_PK per sorted atom: the inputs of T (SpMV) and of the
                                                 // preconditioner, packed with the positions (written by the
                                                 // PCG kernels that produce them)
    // PCG preconditioner: rows keep their pairs within usolvcut first
    bool prec_on = false;                        // this evaluation solves with the sparse preconditioner
    pmq::DBuf<unsigned long long> nmask;         // 128 per tile: the pairs within usolvcut (row, column masks)
    pmq::DBuf<int> nsoff;                        // 128 per tile: offsets within the near part of the row
    pmq::DBuf<int> rownear;                      // npad: near entries per row
    pmq::DBuf<double> tval_s;                    // tcap records: the near pairs' tensors without Ewald
    size_t tcap = 0;
    int nnz = 0;
    bool tmixed = false;                         // tval holds floats (mixed form)
    // batched clusters (k_batch.hip)
    int nsys = 0;                                // 0 = one system
    std::vector<int> h_boff;                     // nsys + 1 atom offsets
    pmq::DBuf<int> b_order, b_off, b_eoff, b_counter;
    pmq::DBuf<double> b_em, b_ep, b_it;
    int num_cus = 256;
    int b_beg[PMQ_BATCH_BUCKETS + 1] = {0, 0, 0, 0}; // size buckets: slices of b_order (>128, >64, <=64 atoms)
    int b_var[2][PMQ_BATCH_BUCKETS] = {{0, 0, 0}, {0, 0, 0}}; // tuned shape per [FP64, mixed][bucket]
    float b_ms[2][PMQ_BATCH_BUCKETS] = {{-1, -1, -1}, {-1, -1, -1}}; // its timed launch
    bool b_tuned[2] = {false, false};
    bool b_tuned_pg[2] = {false, false};         // the parameter-gradient setting the tuning ran with
    bool b_merged[2] = {false, false};           // tuned: all clusters in one 256-row launch
    hipStream_t b_side[PMQ_BATCH_BUCKETS - 1] = {}; // the other buckets' streams (created by pmq_set_batch)
    hipEvent_t b_fork = nullptr, b_join[PMQ_BATCH_BUCKETS - 1] = {};
    // the permanent multipoles' reciprocal potential beside the real-space
    // pair kernels (default priority; created with the PME grids)
    hipStream_t pme_side = nullptr;
    hipEvent_t pme_fork = nullptr, pme_join = nullptr;
    bool pme_forked = false;       // launches went out on pme_side, not joined yet
    bool mpole_frac_ready = false; // pme_mpole_fields ran in this evaluation
    bool disp_pot_ready = false;   // ... and the dispersion potential after it
    bool b_pin[2][PMQ_BATCH_BUCKETS] = {{false, false, false}, {false, false, false}}; // pmq_set_batch_shape
    bool b_pgrad = false;                        // batched GRADIENT evaluations also give parameter gradients
    bool b_pg_valid = false;                     // pgo holds the last batched evaluation's
    pmq::DBuf<double> pgo;                       // 16n: dpole (13n), dpolarity, dpdamp, dthole (caller order)
    // one evaluation of a single system as a HIP graph: captured on the
    // second consecutive speculative evaluation with the same key, replayed
    // while the key holds (pmq_api.cpp eval_key)
    struct EvalKey {
        int mode, precision, pcg_batch;
        unsigned terms;
        unsigned long long epoch;
        pmq::DevBox box;
    };
    unsigned long long epoch = 0;                // bumped by whatever changes launches or their arguments
    bool graphs_on = false;                      // pmq_set_graphs
    EvalKey e_last{}, e_key{};                   // last eager speculative evaluation; the captured one
    bool e_last_ok = false, e_nocapture = false; // e_last is set; capturing e_last failed
    void *e_graph = nullptr, *e_exec = nullptr;
    void *b_graph = nullptr, *b_exec = nullptr;  // captured evaluation
    void *static_scratch = nullptr;              // this device's fixed scratch array (pmq_compute_batch_noalloc)
    bool noalloc_last = false;                   // the last evaluation was pmq_compute_batch_noalloc
    int b_graph_key = -1;
    // PME
    int K[3] = {0, 0, 0};
    int bsorder = 5;
    double pmeA[3][3] = {{0}};                   // u = A r
    pmq::DBuf<double> pme_g;                     // (K0/2+1) K1 K2 reciprocal kernel
    pmq::DBuf<double> pme_gd;                    // the same for dispersion (m = 0 included)
    pmq::DBuf<double> pme_bsmod;                 // B-spline moduli of the three axes (K0, then K1, then K2)
    pmq::DBuf<double> pme_vsrc;                  // VIRIAL: 10 x npad multipoles plus both induced-dipole sets
    double pme_dalpha = 0.0;                     // dispersion Ewald coefficient pme_gd was built for
    pmq::DBuf<double> qgrid;                     // 2 real grids
    pmq::DBuf<double> phi_acc;                   // the solve's carried potential grids: p then u, 2 sets each
    bool phi_carried = false;                    // phi_acc's u grids hold the potential of this solve's dipoles
    pmq::DBuf<double> cgrid;                     // 2 complex half grids (interleaved re, im)
    pmq::DBuf<double> pme_frac;                  // 2 x 20 fractional potential derivatives per atom
    int pme_nb[3] = {1, 1, 1};                   // spreading bricks per axis
    bool pme_cells = false;                      // this evaluation's atom sort ordered atoms by spreading brick
    pmq::DBuf<int> pme_bstart, pme_bend;         // sorted-atom range of each brick
    pmq::DBuf<unsigned char> fft_work;
    void *plan_f1 = nullptr, *plan_b1 = nullptr, *plan_f2 = nullptr, *plan_b2 = nullptr;
    // FP32 plans, kernel table and flag of the mixed form's electrostatic PME
    // (grids in the same buffers, read as float)
    void *plan_f1s = nullptr, *plan_b1s = nullptr, *plan_f2s = nullptr, *plan_b2s = nullptr;
    pmq::DBuf<float> pme_g_s;
    bool pme_fp32 = false;   // FP32 weights and sums in spreading and gathering (mixed form)
    int pme_fft32 = -1;      // mixed form: FP32 grids and FFTs (1) or FP64 (0), timed once per grid (-1)
    int fft_engine = PMQ_FFT_AUTO;
    int fft_grids = 0;                           // mixed-form grid precision: 0 timed, 1 FP32, 2 FP64
    int gather_lanes = PMQ_GATHER_AUTO;
    int pair_form = PMQ_PAIR_CARTESIAN;          // multipole tiles: Cartesian or quasi-internal-frame pair algebra
    int pme_own64 = -1, pme_own32 = -1; // FP64 / FP32 grids: the direct LDS transforms (1) or rocFFT (0), timed once per grid (-1)
    pmq::DBuf<double> fft_tw;    // twiddles of the direct transforms, axes 0..2, FP64 and FP32
    pmq::DBuf<float> fft_tw_s;
    void *fft_info = nullptr;
    pmq::DevBox pme_dbox{};                       // box the kernel G was built for
    double pme_alpha = 0.0;

    int pcg_iters = 0;
    // HIPPO dispersion and repulsion
    pmq::DBuf<double> csix, adisp, sizpr, dmppr, elepr; // caller order
    pmq::DBuf<double> shp;                       // PMQ_SHP_ROWS * npad sorted: csix, adisp, sizpr, dmppr, elepr,
                                                 // chgct, dmpct (0 read as 1000), use[] (1 or 0)
    pmq::DBuf<double> der, dedsp, ter;           // 3n gradients of er, edsp; torques of er
    bool have_disp = false, have_rep = false;
    // van der Waals (k_vdw.hip) and charge transfer
    int nvdw = 0;
    bool have_vdw = false, have_ct = false;
    pmq::DBuf<int> vcls, ired, vuse;             // n: class, reduction partner, reduced site active
    pmq::DBuf<int> vkid_off, vkid;               // n + 1 offsets, list: the atoms reduced toward each atom
    pmq::DBuf<double> kred;                      // n reduction factors
    pmq::DBuf<double> vtab;                      // 4 nvdw^2: radmin, epsilon, radmin4, epsilon4
    pmq::DBuf<double> vx, vy, vz;                // n reduced sites
    pmq::DBuf<double> gvr, dev;                  // 3n: gradient at the reduced sites; of ev at the atoms
    pmq::DBuf<double> chgct, dmpct, dect;        // caller order; 3n gradient of ect
    pmq::DBuf<double> uind_out;                  // 6n caller-order induced dipoles (d then p)

    // kernel timing (pmq_profile_*)
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::map<std::string, std::vector<std::pair<hipEvent_t, hipEvent_t>>> prof;

    int last_mode = PMQ_ENERGY;
    unsigned last_terms = 0;
    int last_precision = PMQ_PLAIN_FP64;
};

// ---------------------------------------------------------------------------
// launchers (defined in the .hip files)
namespace pmq {
// positions and active flags default to the atoms' (x, y, z, usef)
int build_neighbors(pmq_ctx *c, double cutoff, const double *x = nullptr, const double *y = nullptr,
                    const double *z = nullptr, const int *use = nullptr);
// polar_sites: also the sorted polarization parameters and the zeroed
// permanent field; clear_scaled: also the scaled pairs cleared from the tile
// masks (after tile_structure(..., false))
int launch_rotpole(pmq_ctx *c, bool want_rpole13, bool polar_sites = false, bool clear_scaled = false);
int launch_chgflx_delta(pmq_ctx *c);
int launch_mpole_direct(pmq_ctx *c, int mode, double off2, bool mixed);
int launch_fx_flush(pmq_ctx *c, long long *q, double *d, size_t n, double scale);
int launch_fx_flush_pair(pmq_ctx *c, long long *q, double *d0, double *d1, size_t n, double scale);
int launch_torque(pmq_ctx *c, int mode, const double *te, double *de, bool mixed);
int launch_torque_pair(pmq_ctx *c, int mode0, const double *te0, double *de0, int mode1, const double *te1,
                       double *de1, bool mixed);
constexpr int PMQ_ZERO_MAX = 16;
int launch_zero_many(pmq_ctx *c, int k, double *const *p, const size_t *n);
// the status record and PCG scalars into the pinned host mirror (h_status)
int launch_status_out(pmq_ctx *c);
// scalar stride of the per-item tile scalars (tscal)
inline size_t item_stride(const pmq_ctx *c) { return (size_t)PMQ_TILE_PIECES * (size_t)c->tile_cap; }
int pme_setup(pmq_ctx *c);
int fft_own_setup(pmq_ctx *c);
bool fft_own_fits(const pmq_ctx *c, bool fp32);
int fft_own_convolve(pmq_ctx *c, int nset, const void *G, bool fp32, const double *skip = nullptr);
int pme_disp(pmq_ctx *c, bool grad, bool need_bin, bool vir);
int hippo_sort(pmq_ctx *c);
int launch_disp(pmq_ctx *c, double off2, bool grad, bool mixed, bool vir);
int launch_rep(pmq_ctx *c, double off2, bool grad, bool mixed, bool vir);
int launch_ct(pmq_ctx *c, double off2, bool grad, bool mixed, bool vir);
// the van der Waals term with its own pair structure over the reduced
// sites; run before the other terms build theirs
int vdw_pass(pmq_ctx *c, bool grad, bool mixed, bool vir, bool safe);
void pme_release(pmq_ctx *c);
int pme_mpole_fields(pmq_ctx *c, bool side, bool disp = false);
int pme_mpole(pmq_ctx *c, int mode, bool mpole, bool polar);
// VIRIAL: the k-space sums of the multipole and polarization reciprocal
// energies (after the solve; FP64 grids whatever the precision form)
int pme_virial(pmq_ctx *c, bool mpole, bool polar);
int polar_prepare(pmq_ctx *c, double off2, bool mixed, bool safe);
int polar_solve(pmq_ctx *c, bool safe, int *iters_out);
int polar_gradient(pmq_ctx *c, double off2, bool mixed, bool vir);
int polar_finish(pmq_ctx *c, double *ud_out, double *up_out, double *aep = nullptr);
// the charge-flux chain term of the multipole (pot -> dem) and polarization
// (potp -> dep) gradients; mixed: fixed point, flushed
int launch_chgflx_grad(pmq_ctx *c, int mode, bool mpole, bool polar, bool mixed);
int launch_zero(pmq_ctx *c, double *p, size_t n);
// clear_scaled false: the caller clears the scaled pairs from the tile masks
// itself (with the multipole rotation, launch_rotpole)
int tile_structure(pmq_ctx *c, double off2, bool safe, bool clear_scaled = true);
int pair_matrix_build(pmq_ctx *c, bool mixed, bool safe);
// scal_which, k0, k1: the tile-item scalars [k0, k1) summed in the same
// launch (energy component scal_which for scalar 0), or none when k1 <= k0
int tile_reduce(pmq_ctx *c, TileReduceArgs &a, int nv, bool float_parts, bool fixed, int scal_which = 0, int k0 = 0,
                int k1 = 0);
int tile_scalars(pmq_ctx *c, int which, bool full);
int tile_virial(pmq_ctx *c); // tile scalars 1..9 (the pair virial) into the totals
// skip: a device flag; when it is nonzero the launched kernels return at once
// (the speculative PCG iterations after convergence)
// pcg (or null) with mode PCG_APPLY_CHECK: the first launch of a PCG
// iteration (its convergence test first); PCG_APPLY_E1: also the step E1 on
// the rows instead of storing dst (the whole operator when there is no PME
// part), recording the launch's block count in pcg->n1 (pcg_steps.h)
struct PcgVec;
enum { PCG_APPLY_NONE = 0, PCG_APPLY_CHECK = 1, PCG_APPLY_E1 = 2 };
int pair_matrix_apply(pmq_ctx *c, double *dst, const double *skip = nullptr, PcgVec *pcg = nullptr,
                      int mode = PCG_APPLY_NONE);
int pair_matrix_prec_blocks(const pmq_ctx *c);
// E non-null: the solve's first guess a E + max(a, 1e-8) T_s (a E) instead;
// carry: also fold the reciprocal potential grids of z into those of p and u
// (pme_carry_grids)
int pair_matrix_precondition(pmq_ctx *c, const PcgVec &pcg, const double *E = nullptr, bool carry = false);
// the reciprocal potential grid of z (qgrid) and the carried grids of p and u
// with their element type, when PME is set up (false otherwise)
bool pme_carry_grids(pmq_ctx *c, void **qg, void **gp, void **gu, size_t *pts, bool *g32);
// device addresses of the batch scratch when it lives outside the context
struct BatchScratch {
    Totals *tot;
    int *counter;
    double *em_c, *ep_c, *it_c;
};
int launch_batch(pmq_ctx *c, int mode, bool mpole, bool polar, bool mixed, const int *variants,
                 const BatchScratch *bs, bool pgrad, bool merged);
int batch_tune(pmq_ctx *c, int mode, bool mpole, bool polar, bool mixed, bool pgrad);
int batch_variants(int bucket);
int batch_bucket_rows(int bucket);
void batch_shape(int bucket, int variant, int *rows, int *lanes, int *wgs_per_cu);
int launch_param_site(pmq_ctx *c, bool mpole, bool polar, double *dpole, double *dpolarity, double *dpdamp,
                      double *dthole);
int static_scratch_attach(pmq_ctx *c, int device);
size_t static_scratch_bytes();
int batch_noalloc_enqueue(pmq_ctx *c, int mode, bool mpole, bool polar, bool mixed, const int *variants, bool merged,
                          double *em,
                          double *ep, int *iters, double *etot, double *grad);
int param_gradients(pmq_ctx *c, bool mpole, bool polar, double *dpole, double *dpolarity, double *dpdamp,
                    double *dthole);
} // namespace pmq

namespace pmq {
// brackets the launches issued in its lifetime with events when profiling
struct ProfScope {
    pmq_ctx *c;
    const char *name;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(pmq_ctx *ctx, const char *nm);
    ~ProfScope();
};
} // namespace pmq

#define PMQ_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            c->err = std::string(#call) + ": " + hipGetErrorString(e_);                            \
            return PMQ_ERR_HIP;                                                                    \
        }                                                                                          \
    } while (0)
