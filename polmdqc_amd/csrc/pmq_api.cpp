// pmq_api.cpp -- the C ABI of include/pmq.h: context, uploads, evaluation
// sequence and result download.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

#include "pmq_internal.h"

namespace pmq {
size_t neighbor_scratch_bytes(int n, int nblk);
} // namespace pmq

using namespace pmq;

namespace pmq {
static hipEvent_t take_event(pmq_ctx *c)
{
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

ProfScope::ProfScope(pmq_ctx *ctx, const char *nm) : c(ctx), name(nm)
{
    if (!c->profiling) return;
    a = take_event(c);
    b = take_event(c);
    if (a) (void)hipEventRecord(a, c->stream);
}

ProfScope::~ProfScope()
{
    if (!c->profiling || !a || !b) return;
    (void)hipEventRecord(b, c->stream);
    c->prof[name].push_back({a, b});
}
} // namespace pmq

#define CHECK_CTX(c)                                                                               \
    do {                                                                                           \
        if (!(c)) return PMQ_ERR_ARG;                                                              \
    } while (0)

template <class T>
static int upload(pmq_ctx *c, DBuf<T> &d, const T *h, size_t count)
{
    PMQ_HIP(d.alloc(count));
    if (count && h) PMQ_HIP(hipMemcpyAsync(d.p, h, count * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return PMQ_OK;
}

static int pmq_check(pmq_ctx *c);

// Setters first complete an evaluation that is still pending (pmq_check:
// wait, and redo it on its own inputs if it fell short), so a later fetch
// returns the results of the inputs that evaluation was enqueued with.
static void new_epoch(pmq_ctx *c);

#define SETTLE_PENDING(c)                                                                          \
    do {                                                                                           \
        if ((c)->pending) {                                                                        \
            const int rc_ = pmq_check(c);                                                          \
            if (rc_) return rc_;                                                                   \
        }                                                                                          \
    } while (0)

extern "C" {

int pmq_abi_version(void) { return PMQ_ABI_VERSION; }

int pmq_profile_enable(pmq_ctx *c, int on)
{
    CHECK_CTX(c);
    PMQ_HIP(hipStreamSynchronize(c->stream));
    c->profiling = on != 0;
    new_epoch(c); // events are recorded around launches, which a graph does not hold
    c->prof.clear();
    c->ev_used = 0;
    return PMQ_OK;
}

int pmq_profile_read(pmq_ctx *c, const char *name, double *total_ms, long long *launches)
{
    CHECK_CTX(c);
    if (!name || !total_ms || !launches) return PMQ_ERR_ARG;
    PMQ_HIP(hipStreamSynchronize(c->stream));
    double sum = 0.0;
    long long cnt = 0;
    auto it = c->prof.find(name);
    if (it != c->prof.end()) {
        for (auto &pr : it->second) {
            float ms = 0.0f;
            PMQ_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
            sum += ms;
            cnt++;
        }
        it->second.clear();
    }
    bool empty = true;
    for (auto &kv : c->prof) empty = empty && kv.second.empty();
    if (empty) c->ev_used = 0;
    *total_ms = sum;
    *launches = cnt;
    return PMQ_OK;
}

int pmq_set_fft(pmq_ctx *c, int engine)
{
    CHECK_CTX(c);
    const int grids = engine & (PMQ_FFT_GRIDS_FP32 | PMQ_FFT_GRIDS_FP64);
    engine &= ~(PMQ_FFT_GRIDS_FP32 | PMQ_FFT_GRIDS_FP64);
    if (engine < PMQ_FFT_AUTO || engine > PMQ_FFT_DIRECT || grids == (PMQ_FFT_GRIDS_FP32 | PMQ_FFT_GRIDS_FP64)) {
        c->err = "pmq_set_fft: engine must be PMQ_FFT_AUTO, PMQ_FFT_LIBRARY or PMQ_FFT_DIRECT, with at most one of "
                 "PMQ_FFT_GRIDS_FP32 and PMQ_FFT_GRIDS_FP64";
        return PMQ_ERR_ARG;
    }
    SETTLE_PENDING(c);
    new_epoch(c);
    c->fft_engine = engine;
    c->fft_grids = grids == PMQ_FFT_GRIDS_FP32 ? 1 : grids == PMQ_FFT_GRIDS_FP64 ? 2 : 0;
    c->pme_own64 = c->pme_own32 = -1;
    c->pme_fft32 = -1;
    return PMQ_OK;
}

int pmq_set_gather_lanes(pmq_ctx *c, int mode)
{
    CHECK_CTX(c);
    if (mode < PMQ_GATHER_AUTO || mode > PMQ_GATHER_PLANES) {
        c->err = "pmq_set_gather_lanes: mode must be PMQ_GATHER_AUTO, PMQ_GATHER_ROWS or PMQ_GATHER_PLANES";
        return PMQ_ERR_ARG;
    }
    SETTLE_PENDING(c);
    new_epoch(c);
    c->gather_lanes = mode;
    return PMQ_OK;
}

int pmq_set_pair_form(pmq_ctx *c, int form)
{
    CHECK_CTX(c);
    if (form != PMQ_PAIR_CARTESIAN && form != PMQ_PAIR_QI) {
        c->err = "pmq_set_pair_form: form must be PMQ_PAIR_CARTESIAN or PMQ_PAIR_QI";
        return PMQ_ERR_ARG;
    }
    SETTLE_PENDING(c);
    new_epoch(c);
    c->pair_form = form;
    return PMQ_OK;
}

int pmq_fft_choice(pmq_ctx *c, int *fp64_direct, int *fp32_grids, int *fp32_direct)
{
    CHECK_CTX(c);
    if (!fp64_direct || !fp32_grids || !fp32_direct) return PMQ_ERR_ARG;
    *fp64_direct = c->pme_own64;
    *fp32_grids = c->pme_fft32;
    *fp32_direct = c->pme_own32;
    return PMQ_OK;
}

static void eval_graph_release(pmq_ctx *c)
{
    if (c->e_exec) (void)hipGraphExecDestroy((hipGraphExec_t)c->e_exec);
    if (c->e_graph) (void)hipGraphDestroy((hipGraph_t)c->e_graph);
    c->e_exec = c->e_graph = nullptr;
}

// whatever changes the launches of an evaluation or their arguments (sizes,
// buffers, parameters, transform choices, profiling) moves the epoch on, so
// the next evaluation is enqueued afresh
static void new_epoch(pmq_ctx *c)
{
    c->epoch++;
    c->e_last_ok = false;
    c->e_nocapture = false;
}

static pmq_ctx::EvalKey eval_key(const pmq_ctx *c, int mode, unsigned terms, int precision)
{
    pmq_ctx::EvalKey k;
    std::memset(&k, 0, sizeof k); // padding too: keys are compared bytewise
    k.mode = mode;
    k.terms = terms;
    k.precision = precision;
    k.pcg_batch = c->pcg_guess_iters;
    k.epoch = c->epoch;
    k.box = c->box;
    return k;
}

static bool same_key(const pmq_ctx::EvalKey &a, const pmq_ctx::EvalKey &b)
{
    return std::memcmp(&a, &b, sizeof a) == 0;
}

static void batch_graph_release(pmq_ctx *c)
{
    if (c->b_exec) (void)hipGraphExecDestroy((hipGraphExec_t)c->b_exec);
    if (c->b_graph) (void)hipGraphDestroy((hipGraph_t)c->b_graph);
    c->b_exec = c->b_graph = nullptr;
    c->b_graph_key = -1;
}

int pmq_create(const pmq_config *cfg, pmq_ctx **out)
{
    if (!out) return PMQ_ERR_ARG;
    *out = nullptr;
    pmq_ctx *c = new pmq_ctx();
    c->device = cfg ? cfg->device : 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= c->device || c->device < 0) {
        delete c;
        return PMQ_ERR_HIP;
    }
    if (hipSetDevice(c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return PMQ_ERR_HIP;
    }
    if (c->totals.alloc(1) != hipSuccess || c->bbox.alloc(6) != hipSuccess || c->ntiles_d.alloc(2) != hipSuccess ||
        c->ntiles_eff.alloc(1) != hipSuccess || c->status.alloc(1) != hipSuccess ||
        hipHostMalloc((void **)&c->h_status, sizeof(HostStatus), hipHostMallocDefault) != hipSuccess) {
        delete c;
        return PMQ_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess && prop.multiProcessorCount > 0)
        c->num_cus = prop.multiProcessorCount;
    // this device's address of the fixed scratch of pmq_compute_batch_noalloc
    if (static_scratch_attach(c, c->device) != PMQ_OK) {
        pmq_destroy(c);
        return PMQ_ERR_HIP;
    }
    *out = c;
    return PMQ_OK;
}

int pmq_destroy(pmq_ctx *c)
{
    if (!c) return PMQ_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    pme_release(c);
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    eval_graph_release(c);
    batch_graph_release(c);
    for (int k = 0; k < PMQ_BATCH_BUCKETS - 1; k++) {
        if (c->b_side[k]) (void)hipStreamSynchronize(c->b_side[k]);
        if (c->b_side[k]) (void)hipStreamDestroy(c->b_side[k]);
        if (c->b_join[k]) (void)hipEventDestroy(c->b_join[k]);
    }
    if (c->b_fork) (void)hipEventDestroy(c->b_fork);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->h_status) (void)hipHostFree(c->h_status);
    delete c;
    return PMQ_OK;
}

const char *pmq_last_error(const pmq_ctx *c) { return c ? c->err.c_str() : "null context"; }

int pmq_set_topology(pmq_ctx *c, const pmq_topology *t)
{
    CHECK_CTX(c);
    SETTLE_PENDING(c);
    new_epoch(c);
    // a call that fails part way leaves no usable system behind: nothing
    // computes until a later set_topology (and set_params, set_coords) succeeds
    c->have_top = c->have_prm = c->have_xyz = false;
    c->nsys = 0;
    if (!t || t->n <= 0 || !t->polaxe || !t->zaxis || !t->xaxis || !t->yaxis || !t->n12 || !t->n13 || !t->n14 ||
        !t->n15) {
        c->err = "pmq_set_topology: missing arrays or n <= 0";
        return PMQ_ERR_ARG;
    }
    PMQ_HIP(hipSetDevice(c->device));
    const int n = t->n;
    c->n = n;
    c->nblk = (n + PMQ_BLOCK - 1) / PMQ_BLOCK;
    c->npad = c->nblk * PMQ_BLOCK;
    c->h_polaxe.assign(t->polaxe, t->polaxe + n);
    c->h_zaxis.assign(t->zaxis, t->zaxis + n);
    c->h_xaxis.assign(t->xaxis, t->xaxis + n);
    c->h_yaxis.assign(t->yaxis, t->yaxis + n);
    c->h_molcule.assign(n, 0);
    if (t->molcule) c->h_molcule.assign(t->molcule, t->molcule + n);
    for (int i = 0; i < n; i++) {
        const int fr[3] = {t->zaxis[i], t->xaxis[i], std::abs(t->yaxis[i])};
        if (t->polaxe[i] < PMQ_FRAME_NONE || t->polaxe[i] > PMQ_FRAME_THREEFOLD) {
            c->err = "pmq_set_topology: bad frame type";
            return PMQ_ERR_ARG;
        }
        for (int k = 0; k < 3; k++)
            if (fr[k] < 0 || fr[k] > n) {
                c->err = "pmq_set_topology: frame atom out of range";
                return PMQ_ERR_ARG;
            }
        // the frame atoms each frame type reads (rotpole.cpp:102-345, torque.cpp:140-160)
        const int f = t->polaxe[i];
        const int need = f == PMQ_FRAME_NONE ? 0 : f == PMQ_FRAME_ZONLY ? 1
                       : (f == PMQ_FRAME_ZTHENX || f == PMQ_FRAME_BISECTOR) ? 2 : 3;
        if ((need >= 1 && fr[0] == 0) || (need >= 2 && fr[1] == 0) || (need >= 3 && fr[2] == 0)) {
            c->err = "pmq_set_topology: frame type needs a frame atom that is missing";
            return PMQ_ERR_ARG;
        }
    }
    if (t->nbond < 0 || t->nangle < 0 || (t->nbond && !t->ibnd) || (t->nangle && !t->iang)) {
        c->err = "pmq_set_topology: bad bond/angle lists";
        return PMQ_ERR_ARG;
    }
    for (long long j = 0; j < 2LL * t->nbond; j++)
        if (t->ibnd[j] < 0 || t->ibnd[j] >= n) {
            c->err = "pmq_set_topology: bond atom out of range";
            return PMQ_ERR_ARG;
        }
    for (long long j = 0; j < 3LL * t->nangle; j++)
        if (t->iang[j] < 0 || t->iang[j] >= n) {
            c->err = "pmq_set_topology: angle atom out of range";
            return PMQ_ERR_ARG;
        }
    if (t->balist)
        for (long long j = 0; j < 2LL * t->nangle; j++)
            if (t->balist[j] < 0 || t->balist[j] >= std::max(t->nbond, 1)) {
                c->err = "pmq_set_topology: angle bond index out of range";
                return PMQ_ERR_ARG;
            }
    // an atom takes part when it or one of its frame atoms is active (empole.cpp:221,242-243)
    c->h_usef.assign(n, 1);
    c->h_use.assign(n, 1);
    if (t->use)
        for (int i = 0; i < n; i++) c->h_use[i] = t->use[i + 1] ? 1 : 0;
    if (t->use) {
        for (int i = 0; i < n; i++) {
            const int iz = t->zaxis[i], ix = t->xaxis[i], iy = std::abs(t->yaxis[i]);
            c->h_usef[i] = (t->use[i + 1] || t->use[iz] || t->use[ix] || t->use[iy]) ? 1 : 0;
        }
    }
    // scaled pairs i < k.  A later connectivity class overrides an earlier
    // one (empole.cpp:224-235 assigns 1-2, 1-3, 1-4, 1-5 in that order); the
    // polarization-group classes 1-1..1-4 are kept beside them
    c->h_excl.clear();
    {
        const int *cnt[4] = {t->n12, t->n13, t->n14, t->n15};
        const int *lst[4] = {t->i12, t->i13, t->i14, t->i15};
        const int *gcnt[4] = {t->np11, t->np12, t->np13, t->np14};
        const int *glst[4] = {t->ip11, t->ip12, t->ip13, t->ip14};
        std::vector<long long> off[4], goff[4];
        for (int q = 0; q < 4; q++) {
            off[q].assign(n + 1, 0);
            goff[q].assign(n + 1, 0);
            for (int i = 0; i < n; i++) {
                off[q][i + 1] = off[q][i] + cnt[q][i];
                goff[q][i + 1] = goff[q][i] + ((gcnt[q] && glst[q]) ? gcnt[q][i] : 0);
            }
        }
        std::map<int, std::pair<int, int>> cls; // k -> (bonded class, group class)
        std::vector<int> l14;                   // listed 1-4 of i (the vdW 1-4 tables)
        for (int i = 0; i < n; i++) {
            cls.clear();
            l14.clear();
            for (long long j = off[2][i]; j < off[2][i + 1]; j++) l14.push_back(lst[2][j]);
            for (int q = 0; q < 4; q++)
                for (long long j = off[q][i]; j < off[q][i + 1]; j++) {
                    const int k = lst[q][j];
                    if (k < 0 || k >= n) {
                        c->err = "pmq_set_topology: connectivity atom out of range";
                        return PMQ_ERR_ARG;
                    }
                    cls[k].first = q + 2;
                }
            std::vector<int> same;
            for (int q = 0; q < 4; q++)
                for (long long j = goff[q][i]; j < goff[q][i + 1]; j++) {
                    const int k = glst[q][j];
                    if (k < 0 || k >= n) {
                        c->err = "pmq_set_topology: polarization group atom out of range";
                        return PMQ_ERR_ARG;
                    }
                    if (k == i) continue;
                    cls[k].second = q + 1;
                    if (q == 0) same.push_back(k);
                }
            for (auto &kv : cls)
                if (kv.first > i) {
                    ExclPair e{};
                    e.i = i;
                    e.k = kv.first;
                    e.mcls = kv.second.first;
                    e.gcls = kv.second.second;
                    e.intra = std::find(same.begin(), same.end(), kv.first) != same.end() ? 1 : 0;
                    e.in14 = std::find(l14.begin(), l14.end(), kv.first) != l14.end() ? 1 : 0;
                    e.m = e.p = e.d = e.u = 1.0;
                    c->h_excl.push_back(e);
                }
        }
    }
    int rc;
    if ((rc = upload(c, c->polaxe, t->polaxe, n))) return rc;
    if ((rc = upload(c, c->zaxis, t->zaxis, n))) return rc;
    if ((rc = upload(c, c->xaxis, t->xaxis, n))) return rc;
    if ((rc = upload(c, c->yaxis, t->yaxis, n))) return rc;
    if ((rc = upload(c, c->molcule, c->h_molcule.data(), n))) return rc;
    if ((rc = upload(c, c->usef, c->h_usef.data(), n))) return rc;
    if ((rc = upload(c, c->use_raw, c->h_use.data(), n))) return rc;
    c->nbond = t->nbond;
    c->nangle = t->nangle;
    if ((rc = upload(c, c->ibnd, t->ibnd, 2 * (size_t)t->nbond))) return rc;
    if ((rc = upload(c, c->iang, t->iang, 3 * (size_t)t->nangle))) return rc;
    if ((rc = upload(c, c->balist, t->balist, t->balist ? 2 * (size_t)t->nangle : 0))) return rc;

    // per-eval buffers sized for this topology
    const int npad = c->npad, nblk = c->nblk;
    const long long npairs = (long long)nblk * (nblk + 1) / 2;
    PMQ_HIP(c->x.alloc(n)); PMQ_HIP(c->y.alloc(n)); PMQ_HIP(c->z.alloc(n));
    PMQ_HIP(c->pdelta.alloc(n));
    PMQ_HIP(c->potp.alloc(n));
    PMQ_HIP(c->fx_pot.alloc(2 * (size_t)n));
    PMQ_HIP(c->rpole13.alloc(13 * (size_t)n));
    PMQ_HIP(c->perm.alloc(npad)); PMQ_HIP(c->iperm.alloc(n));
    PMQ_HIP(c->sx.alloc(npad)); PMQ_HIP(c->sy.alloc(npad)); PMQ_HIP(c->sz.alloc(npad));
    PMQ_HIP(c->smp.alloc(10 * (size_t)npad)); PMQ_HIP(c->scp.alloc(3 * (size_t)npad));
    PMQ_HIP(c->sflag.alloc(npad)); PMQ_HIP(c->smol.alloc(npad));
    PMQ_HIP(c->keys.alloc(n)); PMQ_HIP(c->keys_sorted.alloc(n));
    PMQ_HIP(c->blk.alloc(PMQ_BLK_STRIDE * (size_t)nblk));
    PMQ_HIP(c->tileflag.alloc(npairs)); PMQ_HIP(c->tilescan.alloc(npairs));
    PMQ_HIP(c->tiles.alloc(npairs));
    c->ntiles_cap = (int)std::min<long long>(npairs, 1 << 30);
    c->scratch_sort_bytes = neighbor_scratch_bytes(n, nblk);
    PMQ_HIP(c->scratch_sort.alloc(c->scratch_sort_bytes));
    PMQ_HIP(c->dem.alloc(3 * (size_t)n)); PMQ_HIP(c->te.alloc(3 * (size_t)n));
    PMQ_HIP(c->pot.alloc(n)); PMQ_HIP(c->aem.alloc(n)); PMQ_HIP(c->aep.alloc(n));
    PMQ_HIP(c->rotm.alloc(9 * (size_t)n)); PMQ_HIP(c->flip.alloc(n));
    PMQ_HIP(c->cphi.alloc(10 * (size_t)npad));
    // fixed-point accumulators start (and are left) at zero
    PMQ_HIP(c->fx_grad.alloc(6 * (size_t)n)); PMQ_HIP(c->fx_field.alloc(6 * (size_t)npad));
    int rz;
    if ((rz = launch_zero(c, (double *)c->fx_grad.p, 6 * (size_t)n))) return rz;
    if ((rz = launch_zero(c, (double *)c->fx_field.p, 6 * (size_t)npad))) return rz;
    PMQ_HIP(hipStreamSynchronize(c->stream));
    c->have_top = true;
    c->have_prm = false;
    c->have_xyz = false;
    c->tile_cap = 0;
    c->tcap = 0;
    c->nsys = 0;
    batch_graph_release(c);
    c->safe = true;
    c->pending = false;
    c->pcg_guess_iters = 0;
    return PMQ_OK;
}

int pmq_set_params(pmq_ctx *c, const pmq_params *p)
{
    CHECK_CTX(c);
    SETTLE_PENDING(c);
    new_epoch(c);
    // a rejected call leaves no parameters set (the device copies may be
    // half updated): compute fails with PMQ_ERR_STATE until a call succeeds
    c->have_prm = false;
    batch_graph_release(c);
    if (!c->have_top) {
        c->err = "pmq_set_params: call pmq_set_topology first";
        return PMQ_ERR_STATE;
    }
    if (!p || !p->pole) {
        c->err = "pmq_set_params: pole is required";
        return PMQ_ERR_ARG;
    }
    if (p->pentype < PMQ_PEN_NONE || p->pentype > PMQ_PEN_GORDON2) {
        c->err = "pmq_set_params: bad pentype";
        return PMQ_ERR_ARG;
    }
    if (p->pentype != PMQ_PEN_NONE && (!p->pcore || !p->pval0 || !p->palpha)) {
        c->err = "pmq_set_params: charge penetration needs pcore, pval0, palpha";
        return PMQ_ERR_ARG;
    }
    if (p->use_chgflx && (c->nbond + c->nangle > 0) &&
        (!p->bl || !p->bflx || (c->nangle && (!p->anat || !p->aflx || !p->abflx)))) {
        c->err = "pmq_set_params: charge flux needs bl, bflx, anat, aflx, abflx";
        return PMQ_ERR_ARG;
    }
    if (p->use_chgflx && c->nangle > 0 && c->balist.n < 2 * (size_t)c->nangle) {
        c->err = "pmq_set_params: charge flux over angles needs the topology's balist";
        return PMQ_ERR_ARG;
    }
    if (p->electric == 0.0 || p->dielec == 0.0) {
        c->err = "pmq_set_params: electric and dielec must be nonzero";
        return PMQ_ERR_ARG;
    }
    if (p->pcgprec && !(p->usolvcut >= 0.0 && std::isfinite(p->udiag))) {
        c->err = "pmq_set_params: the PCG preconditioner needs usolvcut >= 0 and a finite udiag";
        return PMQ_ERR_ARG;
    }
    PMQ_HIP(hipSetDevice(c->device));
    const int n = c->n;
    int rc;
    if ((rc = upload(c, c->pole, p->pole, 13 * (size_t)n))) return rc;
    const bool cp = p->pentype != PMQ_PEN_NONE;
    if ((rc = upload(c, c->pcore, cp ? p->pcore : nullptr, n))) return rc;
    if ((rc = upload(c, c->pval0, cp ? p->pval0 : nullptr, n))) return rc;
    if ((rc = upload(c, c->palpha, cp ? p->palpha : nullptr, n))) return rc;
    if (p->use_chgflx) {
        if ((rc = upload(c, c->bl, p->bl, c->nbond))) return rc;
        if ((rc = upload(c, c->bflx, p->bflx, c->nbond))) return rc;
        if ((rc = upload(c, c->anat, p->anat, c->nangle))) return rc;
        if ((rc = upload(c, c->aflx, p->aflx, 2 * (size_t)c->nangle))) return rc;
        if ((rc = upload(c, c->abflx, p->abflx, 2 * (size_t)c->nangle))) return rc;
    }
    // scale values of the scaled pairs (empole.cpp:224-235 for m; p, d and u
    // as polpot.h:47-66 names them: p by connectivity with the "intra" value
    // inside one polarization group, d and u by group connectivity)
    for (auto &e : c->h_excl) {
        const double mv[6] = {1, 1, p->m2scale, p->m3scale, p->m4scale, p->m5scale};
        const double pv[6] = {1, 1, p->p2scale, p->p3scale, p->p4scale, p->p5scale};
        const double piv[6] = {1, 1, p->p2iscale, p->p3iscale, p->p4iscale, p->p5iscale};
        const double dv[5] = {1, p->d1scale, p->d2scale, p->d3scale, p->d4scale};
        const double uv[5] = {1, p->u1scale, p->u2scale, p->u3scale, p->u4scale};
        const double wv[6] = {1, 1, p->w2scale, p->w3scale, p->w4scale, p->w5scale};
        const double dsv[6] = {1, 1, p->dsp2scale, p->dsp3scale, p->dsp4scale, p->dsp5scale};
        const double rv[6] = {1, 1, p->r2scale, p->r3scale, p->r4scale, p->r5scale};
        e.m = mv[e.mcls];
        e.p = e.intra ? piv[e.mcls] : pv[e.mcls];
        e.d = p->dpequal ? e.p : dv[e.gcls];
        // the dipole-dipole scale: u by group (AMOEBA), w by connectivity
        // (HIPPO, i.e. with charge penetration)
        e.u = p->pentype != PMQ_PEN_NONE ? wv[e.mcls] : uv[e.gcls];
        e.dsp = dsv[e.mcls];
        e.rep = rv[e.mcls];
        const double vv[6] = {1, 1, p->v2scale, p->v3scale, p->v4scale, p->v5scale};
        e.v = vv[e.mcls];
    }
    if ((rc = upload(c, c->excl, c->h_excl.data(), c->h_excl.size()))) return rc;
    const bool polar = p->polarity != nullptr;
    if (polar) {
        if (!p->thole || !p->pdamp) {
            c->err = "pmq_set_params: polarization needs polarity, thole and pdamp";
            return PMQ_ERR_ARG;
        }
        if ((rc = upload(c, c->polarity, p->polarity, n))) return rc;
        if ((rc = upload(c, c->thole, p->thole, n))) return rc;
        if ((rc = upload(c, c->pdamp, p->pdamp, n))) return rc;
        c->h_polarity.assign(p->polarity, p->polarity + n);
        for (double v : c->h_polarity)
            if (!(v >= 0.0)) {
                c->err = "pmq_set_params: negative polarizability";
                return PMQ_ERR_ARG;
            }
        const size_t nv = 6 * (size_t)c->npad;
        PMQ_HIP(c->spol.alloc(3 * (size_t)c->npad));
        PMQ_HIP(c->cphi.alloc(10 * (size_t)c->npad));
        PMQ_HIP(c->pfield.alloc(nv)); PMQ_HIP(c->uind.alloc(nv));
        PMQ_HIP(c->pres.alloc(nv)); PMQ_HIP(c->pz.alloc(nv));
        PMQ_HIP(c->pdir.alloc(nv)); PMQ_HIP(c->pvec.alloc(nv));
        PMQ_HIP(c->pav.alloc(nv));
        PMQ_HIP(c->pcg.alloc(16));
        PMQ_HIP(c->tep.alloc(3 * (size_t)n)); PMQ_HIP(c->dep.alloc(3 * (size_t)n));
        PMQ_HIP(c->uind_out.alloc(6 * (size_t)n));
    } else {
        c->h_polarity.clear();
    }
    if (p->use_ewald && !(p->aewald > 0.0 && p->ewaldcut > 0.0)) {
        c->err = "pmq_set_params: Ewald needs aewald > 0 and ewaldcut > 0";
        return PMQ_ERR_ARG;
    }
    // HIPPO dispersion and repulsion
    c->have_disp = p->csix != nullptr;
    c->have_rep = p->sizpr != nullptr;
    if (c->have_disp) {
        if (!p->adisp || !(p->dispcut > 0.0)) {
            c->err = "pmq_set_params: dispersion needs csix, adisp and dispcut > 0";
            return PMQ_ERR_ARG;
        }
        if (p->use_dewald && !(p->dewald > 0.0)) {
            c->err = "pmq_set_params: dispersion Ewald needs dewald > 0";
            return PMQ_ERR_ARG;
        }
        if ((rc = upload(c, c->csix, p->csix, n))) return rc;
        if ((rc = upload(c, c->adisp, p->adisp, n))) return rc;
        for (int i = 0; i < n; i++)
            if (p->csix[i] != 0.0 && !(p->adisp[i] > 0.0)) {
                c->err = "pmq_set_params: dispersion damping exponents must be positive";
                return PMQ_ERR_ARG;
            }
    }
    if (c->have_rep) {
        if (!p->dmppr || !p->elepr || !(p->repcut > 0.0)) {
            c->err = "pmq_set_params: repulsion needs sizpr, dmppr, elepr and repcut > 0";
            return PMQ_ERR_ARG;
        }
        for (int i = 0; i < n; i++)
            if (p->sizpr[i] != 0.0 && !(p->dmppr[i] > 0.0)) {
                c->err = "pmq_set_params: repulsion damping exponents must be positive";
                return PMQ_ERR_ARG;
            }
        if ((rc = upload(c, c->sizpr, p->sizpr, n))) return rc;
        if ((rc = upload(c, c->dmppr, p->dmppr, n))) return rc;
        if ((rc = upload(c, c->elepr, p->elepr, n))) return rc;
    }
    // van der Waals
    c->have_vdw = p->nvdw > 0;
    if (c->have_vdw) {
        const int nv = p->nvdw;
        if (!p->vdwclass || !p->radmin || !p->epsilon || !p->radmin4 || !p->epsilon4 || !p->kred || !p->ired ||
            !(p->vdwcut > 0.0)) {
            c->err = "pmq_set_params: van der Waals needs vdwclass, the four tables, kred, ired and vdwcut > 0";
            return PMQ_ERR_ARG;
        }
        if (nv > 4096) {
            c->err = "pmq_set_params: at most 4096 van der Waals classes";
            return PMQ_ERR_ARG;
        }
        std::vector<int> vuse(n), koff(n + 1, 0), kid;
        for (int i = 0; i < n; i++) {
            if (p->vdwclass[i] < 0 || p->vdwclass[i] >= nv || p->ired[i] < 0 || p->ired[i] >= n ||
                !std::isfinite(p->kred[i])) {
                c->err = "pmq_set_params: van der Waals class, reduction partner or factor out of range";
                return PMQ_ERR_ARG;
            }
            // a reduced site is active when its atom or the atom's partner is
            vuse[i] = c->h_use[i] || c->h_use[p->ired[i]];
            if (p->ired[i] != i) koff[p->ired[i] + 1]++;
        }
        for (int i = 0; i < n; i++) koff[i + 1] += koff[i];
        kid.assign(std::max(1, koff[n]), 0);
        {
            std::vector<int> fill(koff.begin(), koff.end() - 1);
            for (int j = 0; j < n; j++)
                if (p->ired[j] != j) kid[fill[p->ired[j]]++] = j; // ascending j within each list
        }
        std::vector<double> tab(4 * (size_t)nv * nv);
        const double *src[4] = {p->radmin, p->epsilon, p->radmin4, p->epsilon4};
        for (int q = 0; q < 4; q++) std::copy(src[q], src[q] + (size_t)nv * nv, tab.begin() + (size_t)q * nv * nv);
        if ((rc = upload(c, c->vcls, p->vdwclass, n))) return rc;
        if ((rc = upload(c, c->ired, p->ired, n))) return rc;
        if ((rc = upload(c, c->kred, p->kred, n))) return rc;
        if ((rc = upload(c, c->vuse, vuse.data(), n))) return rc;
        if ((rc = upload(c, c->vkid_off, koff.data(), n + 1))) return rc;
        if ((rc = upload(c, c->vkid, kid.data(), kid.size()))) return rc;
        if ((rc = upload(c, c->vtab, tab.data(), tab.size()))) return rc;
        PMQ_HIP(c->vx.alloc(n)); PMQ_HIP(c->vy.alloc(n)); PMQ_HIP(c->vz.alloc(n));
        PMQ_HIP(c->gvr.alloc(3 * (size_t)n)); PMQ_HIP(c->dev.alloc(3 * (size_t)n));
        c->nvdw = nv;
        PMQ_HIP(hipStreamSynchronize(c->stream)); // the copies read these host vectors
    }
    // charge transfer
    c->have_ct = p->chgct != nullptr;
    if (c->have_ct) {
        if (!p->dmpct || !(p->ctrncut > 0.0)) {
            c->err = "pmq_set_params: charge transfer needs chgct, dmpct and ctrncut > 0";
            return PMQ_ERR_ARG;
        }
        for (int i = 0; i < n; i++)
            if (!(p->dmpct[i] >= 0.0) || !std::isfinite(p->chgct[i])) {
                c->err = "pmq_set_params: charge-transfer damping exponents must be >= 0";
                return PMQ_ERR_ARG;
            }
        if ((rc = upload(c, c->chgct, p->chgct, n))) return rc;
        if ((rc = upload(c, c->dmpct, p->dmpct, n))) return rc;
        PMQ_HIP(c->dect.alloc(3 * (size_t)n));
    }
    if (c->have_disp || c->have_rep || c->have_ct) {
        PMQ_HIP(c->shp.alloc(PMQ_SHP_ROWS * (size_t)c->npad));
        PMQ_HIP(c->der.alloc(3 * (size_t)n));
        PMQ_HIP(c->dedsp.alloc(3 * (size_t)n));
        PMQ_HIP(c->ter.alloc(3 * (size_t)n));
    }
    c->prm = *p;
    c->prm.pole = nullptr;
    c->prm.pcore = c->prm.pval0 = c->prm.palpha = nullptr;
    c->prm.bl = c->prm.bflx = c->prm.anat = c->prm.aflx = c->prm.abflx = nullptr;
    c->prm.polarity = c->prm.thole = c->prm.pdamp = nullptr;
    c->prm.csix = c->prm.adisp = c->prm.sizpr = c->prm.dmppr = c->prm.elepr = nullptr;
    c->prm.vdwclass = c->prm.ired = nullptr;
    c->prm.radmin = c->prm.epsilon = c->prm.radmin4 = c->prm.epsilon4 = c->prm.kred = nullptr;
    c->prm.chgct = c->prm.dmpct = nullptr;
    PMQ_HIP(hipStreamSynchronize(c->stream));
    c->have_prm = true;
    batch_graph_release(c);
    c->b_tuned[0] = c->b_tuned[1] = false;
    c->b_pg_valid = false;
    c->safe = true;
    c->pending = false;
    c->pcg_guess_iters = 0;
    return PMQ_OK;
}

int pmq_set_coords(pmq_ctx *c, const double *x, const double *y, const double *z, const pmq_box *box)
{
    CHECK_CTX(c);
    SETTLE_PENDING(c);
    if (!c->have_top) {
        c->err = "pmq_set_coords: call pmq_set_topology first";
        return PMQ_ERR_STATE;
    }
    if (!x || !y || !z) {
        c->err = "pmq_set_coords: null coordinates";
        return PMQ_ERR_ARG;
    }
    if (box && (box->kind < PMQ_BOX_NONE || box->kind > PMQ_BOX_DODECA)) {
        c->err = "pmq_set_coords: bad box kind";
        return PMQ_ERR_ARG;
    }
    if (box && box->kind != PMQ_BOX_NONE && !(box->xbox > 0 && box->ybox > 0 && box->zbox > 0)) {
        c->err = "pmq_set_coords: box lengths must be positive";
        return PMQ_ERR_ARG;
    }
    PMQ_HIP(hipSetDevice(c->device));
    const size_t bytes = c->n * sizeof(double);
    PMQ_HIP(hipMemcpyAsync(c->x.p, x, bytes, hipMemcpyHostToDevice, c->stream));
    PMQ_HIP(hipMemcpyAsync(c->y.p, y, bytes, hipMemcpyHostToDevice, c->stream));
    PMQ_HIP(hipMemcpyAsync(c->z.p, z, bytes, hipMemcpyHostToDevice, c->stream));
    c->box = make_box(box);
    PMQ_HIP(hipStreamSynchronize(c->stream));
    c->have_xyz = true;
    return PMQ_OK;
}

static int check_compute(pmq_ctx *c, int mode, unsigned terms, int precision)
{
    if (!c->have_top || !c->have_prm || !c->have_xyz) {
        c->err = "pmq_compute: topology, parameters and coordinates must be set first";
        return PMQ_ERR_STATE;
    }
    if (mode < PMQ_ENERGY || mode > PMQ_VIRIAL) {
        c->err = "pmq_compute: bad mode";
        return PMQ_ERR_ARG;
    }
    if (precision != PMQ_PLAIN_FP64 && precision != PMQ_MIXED) {
        c->err = "pmq_compute: precision must be PMQ_PLAIN_FP64 or PMQ_MIXED";
        return PMQ_ERR_ARG;
    }
    const bool mixed = precision == PMQ_MIXED;
    // charge flux moves only the monopoles, which the multipole and
    // polarization terms see
    if (terms & ~(unsigned)(PMQ_TERM_MPOLE | PMQ_TERM_POLAR | PMQ_TERM_REPEL | PMQ_TERM_DISP | PMQ_TERM_VDW |
                            PMQ_TERM_CTRN)) {
        c->err = "pmq_compute: unknown term";
        return PMQ_ERR_ARG;
    }
    const bool do_vdw = terms & PMQ_TERM_VDW, do_ct = terms & PMQ_TERM_CTRN;
    if ((do_vdw && !c->have_vdw) || (do_ct && !c->have_ct)) {
        c->err = "pmq_compute: van der Waals or charge transfer requested but its parameters were not set";
        return PMQ_ERR_STATE;
    }
    if ((do_vdw || do_ct) && mode == PMQ_ANALYSIS) {
        c->err = "pmq_compute: van der Waals and charge transfer compute energy, gradient and virial (no analysis)";
        return PMQ_ERR_UNSUPPORTED;
    }
    const bool do_mpole = terms & PMQ_TERM_MPOLE, do_polar = terms & PMQ_TERM_POLAR;
    const bool do_rep = terms & PMQ_TERM_REPEL, do_disp = terms & PMQ_TERM_DISP;
    if ((do_rep && !c->have_rep) || (do_disp && !c->have_disp)) {
        c->err = "pmq_compute: repulsion or dispersion requested but its parameters were not set";
        return PMQ_ERR_STATE;
    }
    if ((do_rep || do_disp) && mode == PMQ_ANALYSIS) {
        c->err = "pmq_compute: repulsion and dispersion compute energy, gradient and virial (no analysis)";
        return PMQ_ERR_UNSUPPORTED;
    }
    if (do_rep || do_disp || do_ct) {
        // the pair structure is built at the electrostatic cutoff when those
        // terms run; repcut, dispcut and ctrncut must fit inside it
        const double ecut = c->prm.use_ewald ? c->prm.ewaldcut : c->prm.mpolecut;
        if ((do_mpole || do_polar) && c->box.kind != PMQ_BOX_NONE &&
            ((do_rep && c->prm.repcut > ecut) || (do_disp && c->prm.dispcut > ecut) ||
             (do_ct && c->prm.ctrncut > ecut))) {
            c->err = "pmq_compute: repcut, dispcut and ctrncut may not exceed the electrostatic cutoff";
            return PMQ_ERR_UNSUPPORTED;
        }
    }
    const bool ewald = c->prm.use_ewald;
    if (do_polar && c->h_polarity.empty()) {
        c->err = "pmq_compute: polarization requested but no polarizabilities were set";
        return PMQ_ERR_STATE;
    }
    if (do_polar && c->prm.pentype == PMQ_PEN_GORDON2) {
        c->err = "pmq_compute: HIPPO polarization damping is built for the Gordon1 charge penetration";
        return PMQ_ERR_UNSUPPORTED;
    }
    return PMQ_OK;
}

static int compute_enqueue_body(pmq_ctx *c, int mode, unsigned terms, int precision, bool safe);

// enqueue one evaluation; `safe` sizes the tile and pair-matrix buffers and
// runs the PCG with host round trips, otherwise nothing waits for the device
// and the run is checked afterwards (pmq_check).  Launches forked onto the
// PME side stream are joined back into the context's stream also when the
// evaluation stops early, so that waiting on that stream covers them.
static int compute_enqueue(pmq_ctx *c, int mode, unsigned terms, int precision, bool safe)
{
    const int rc = compute_enqueue_body(c, mode, terms, precision, safe);
    if (c->pme_forked) {
        c->pme_forked = false;
        const hipError_t e = hipStreamWaitEvent(c->stream, c->pme_join, 0);
        if (!rc && e != hipSuccess) {
            c->err = hipGetErrorString(e);
            return PMQ_ERR_HIP;
        }
    }
    return rc;
}

static int compute_enqueue_body(pmq_ctx *c, int mode, unsigned terms, int precision, bool safe)
{
    const bool mixed = precision == PMQ_MIXED;
    const bool do_mpole = terms & PMQ_TERM_MPOLE, do_polar = terms & PMQ_TERM_POLAR;
    const bool do_rep = terms & PMQ_TERM_REPEL, do_disp = terms & PMQ_TERM_DISP;
    const bool do_vdw = terms & PMQ_TERM_VDW, do_ct = terms & PMQ_TERM_CTRN;
    const bool ewald = c->prm.use_ewald;
    const bool dewald = do_disp && c->prm.use_dewald && c->box.kind != PMQ_BOX_NONE;
    PMQ_HIP(hipSetDevice(c->device));
    const int n = c->n;
    int rc;
    const bool grad = mode == PMQ_GRADIENT || mode == PMQ_VIRIAL;
    const bool vir = mode == PMQ_VIRIAL;
    const int fmode = vir ? PMQ_VIRIAL : PMQ_GRADIENT; // the mode of the torque (frame) conversions
    {
        // the evaluation's accumulators, cleared by one launch
        double *zp[PMQ_ZERO_MAX];
        size_t zn[PMQ_ZERO_MAX];
        int k = 0;
        auto add = [&](double *p, size_t cnt) { zp[k] = p; zn[k] = cnt; k++; };
        add((double *)c->totals.p, sizeof(Totals) / sizeof(double));
        static_assert(sizeof(RunStatus) % sizeof(double) == 0, "RunStatus is cleared as doubles");
        add((double *)c->status.p, sizeof(RunStatus) / sizeof(double)); // all-zero bits: every field 0
        if (grad) {
            add(c->dem.p, 3 * (size_t)n);
            add(c->te.p, 3 * (size_t)n);
            add(c->pot.p, n);
            if (do_polar) {
                add(c->dep.p, 3 * (size_t)n);
                add(c->tep.p, 3 * (size_t)n);
            }
            if (c->prm.use_chgflx) {
                add(c->potp.p, n);
                if (mixed) add((double *)c->fx_pot.p, 2 * (size_t)n); // all-zero bits: 0
            }
            if (do_rep) {
                add(c->der.p, 3 * (size_t)n);
                add(c->ter.p, 3 * (size_t)n);
            }
            if (do_disp) add(c->dedsp.p, 3 * (size_t)n);
            if (do_vdw) add(c->gvr.p, 3 * (size_t)n);
            if (do_ct) add(c->dect.p, 3 * (size_t)n);
        }
        if (mode == PMQ_ANALYSIS) add(c->aem.p, n);
        if ((rc = launch_zero_many(c, k, zp, zn))) return rc;
    }
    c->pcg_iters = 0;
    // van der Waals first: its structure over the reduced sites is replaced
    // by the one the other terms build (which pmq_param_grad reuses)
    if (do_vdw && (rc = vdw_pass(c, grad, mixed, vir, safe))) return rc;
    if (do_mpole || do_polar || do_rep || do_disp || do_ct) {
        double cut = ewald ? c->prm.ewaldcut : c->prm.mpolecut;
        const double hcut = std::max({do_rep ? c->prm.repcut : 0.0, do_disp ? c->prm.dispcut : 0.0,
                                      do_ct ? c->prm.ctrncut : 0.0});
        if (!(do_mpole || do_polar)) cut = hcut;
        if (c->box.kind == PMQ_BOX_NONE) cut = std::min(std::max(cut, hcut), 1e12);
        const double off2 = cut * cut;
        const bool cf = c->prm.pentype != PMQ_PEN_NONE && c->prm.use_chgflx;
        if (c->prm.use_chgflx && (rc = launch_chgflx_delta(c))) return rc;
        c->prec_on = do_polar && c->prm.pcgprec && c->prm.usolvcut > 0.0;
        // with PME the atom sort follows the grid's spreading bricks
        c->pme_fp32 = mixed; // electrostatic grids in FP32 for the mixed form
        const bool pme = (ewald && (do_mpole || do_polar)) || dewald;
        if (pme && (rc = pme_setup(c))) return rc;
        c->pme_cells = pme;
        c->mpole_frac_ready = c->disp_pot_ready = false;
        {
            ProfScope ps(c, "neighbors");
            if ((rc = build_neighbors(c, cut))) return rc;
            if ((rc = tile_structure(c, off2, safe, false))) return rc;
        }
        // with the rotation: the sorted polarization parameters, the zeroed
        // permanent field and the scaled pairs cleared from the tile masks
        if ((rc = launch_rotpole(c, true, do_polar, true))) return rc;
        if ((do_rep || do_disp || do_ct) && (rc = hippo_sort(c))) return rc;
        // the permanent reciprocal potential on a second stream, joined
        // before its energy and fields are taken (pme_mpole)
        if (pme && ewald && (do_mpole || do_polar) && (rc = pme_mpole_fields(c, true, dewald))) return rc;
        if (do_rep) {
            const double rc2 = std::min(c->prm.repcut * c->prm.repcut, off2);
            if ((rc = launch_rep(c, rc2, grad, mixed, vir))) return rc;
            if (grad && (rc = launch_torque(c, fmode, c->ter.p, c->der.p, mixed))) return rc;
        }
        if (do_disp) {
            const double dc2 = std::min(c->prm.dispcut * c->prm.dispcut, off2);
            if ((rc = launch_disp(c, dc2, grad, mixed, vir))) return rc;
        }
        if (do_ct) {
            const double cc2 = std::min(c->prm.ctrncut * c->prm.ctrncut, off2);
            if ((rc = launch_ct(c, cc2, grad, mixed, vir))) return rc;
        }
        if (do_polar && (rc = polar_prepare(c, off2, mixed, safe))) return rc;
        if (do_mpole && (rc = launch_mpole_direct(c, mode, off2, mixed))) return rc;
        if (pme) {
            if (ewald && (do_mpole || do_polar) && (rc = pme_mpole(c, mode, do_mpole, do_polar))) return rc;
            if (dewald && (rc = pme_disp(c, grad, !(ewald && (do_mpole || do_polar)), vir))) return rc;
        }
        if (do_polar) {
            if ((rc = polar_solve(c, safe, &c->pcg_iters))) return rc;
            if (grad && (rc = polar_gradient(c, off2, mixed, vir))) return rc;
            if ((rc = polar_finish(c, c->uind_out.p, c->uind_out.p + 3 * (size_t)n,
                                   mode == PMQ_ANALYSIS ? c->aep.p : nullptr)))
                return rc;
        }
        // the k-space virial sums of the reciprocal energies (after the solve)
        if (vir && ewald && (do_mpole || do_polar) && (rc = pme_virial(c, do_mpole, do_polar))) return rc;
        // the torques of both terms in one launch (each adds its frame virial
        // in VIRIAL mode)
        if (grad && do_mpole && do_polar) {
            if ((rc = launch_torque_pair(c, mode, c->te.p, c->dem.p, fmode, c->tep.p, c->dep.p, mixed)))
                return rc;
        } else if (grad && do_polar) {
            if ((rc = launch_torque(c, fmode, c->tep.p, c->dep.p, mixed))) return rc;
        } else if (grad && do_mpole) {
            if ((rc = launch_torque(c, mode, c->te.p, c->dem.p, mixed))) return rc;
        }
        // charge flux: the chain term of both terms' gradients through the
        // monopole shifts (dcflux.cpp:31-225), from the dE/dq of each
        if (grad && cf && (do_mpole || do_polar) && (rc = launch_chgflx_grad(c, mode, do_mpole, do_polar, mixed)))
            return rc;
    }
    c->last_mode = mode;
    c->last_terms = terms;
    c->last_precision = precision;
    // what the check needs, written to the pinned host mirror as the
    // evaluation's last step
    return launch_status_out(c);
}

// one evaluation of the batched clusters: frames, the persistent cluster
// kernels (one per size bucket), torques, and with parameter gradients the
// per-site step that turns them into local-frame derivatives
static int batch_enqueue(pmq_ctx *c, int mode, unsigned terms, int precision, const int *variants)
{
    const bool mixed = precision == PMQ_MIXED;
    const bool mp = terms & PMQ_TERM_MPOLE, pol = terms & PMQ_TERM_POLAR;
    const bool pg = c->b_pgrad && mode == PMQ_GRADIENT;
    int rc;
    if ((rc = launch_zero(c, (double *)c->totals.p, sizeof(Totals) / sizeof(double)))) return rc;
    if ((rc = launch_rotpole(c, false))) return rc;
    const bool merged = c->b_merged[mixed ? 1 : 0];
    if ((rc = launch_batch(c, mode, mp, pol, mixed, variants, nullptr, pg, merged))) return rc;
    if (mode == PMQ_GRADIENT) {
        if (mp && (rc = launch_torque(c, PMQ_GRADIENT, c->te.p, c->dem.p, mixed))) return rc;
        if (pol && (rc = launch_torque(c, PMQ_GRADIENT, c->tep.p, c->dep.p, mixed))) return rc;
    }
    if (pg) {
        const size_t n = c->n;
        double *o = c->pgo.p;
        if ((rc = launch_param_site(c, mp, pol, o, o + 13 * n, o + 14 * n, o + 15 * n))) return rc;
    }
    return PMQ_OK;
}

static int batch_check(pmq_ctx *c, int mode, unsigned terms)
{
    if (mode != PMQ_ENERGY && mode != PMQ_GRADIENT) {
        c->err = "pmq_compute: batched clusters compute energy and gradient";
        return PMQ_ERR_UNSUPPORTED;
    }
    if (terms & (PMQ_TERM_REPEL | PMQ_TERM_DISP | PMQ_TERM_VDW | PMQ_TERM_CTRN)) {
        c->err = "pmq_compute: batched clusters compute the multipole and polarization terms";
        return PMQ_ERR_UNSUPPORTED;
    }
    if (c->prm.pentype != PMQ_PEN_NONE || c->prm.use_chgflx || c->prm.use_ewald || c->box.kind != PMQ_BOX_NONE) {
        c->err = "pmq_compute: batched clusters take non-periodic systems without Ewald, charge penetration or flux";
        return PMQ_ERR_UNSUPPORTED;
    }
    if ((terms & PMQ_TERM_POLAR) && c->h_polarity.empty()) {
        c->err = "pmq_compute: polarization requested but no polarizabilities were set";
        return PMQ_ERR_STATE;
    }
    return PMQ_OK;
}

// buffers of the batched parameter gradients; true if any moved (a captured
// graph holds their addresses)
static int batch_pgrad_buffers(pmq_ctx *c, bool *moved)
{
    const void *before[5] = {c->pg_dm.p, c->pg_dpd.p, c->pg_dth.p, c->pgo.p, c->uind.p};
    PMQ_HIP(c->pg_dm.alloc(10 * (size_t)c->npad));
    PMQ_HIP(c->pg_dpd.alloc(c->npad));
    PMQ_HIP(c->pg_dth.alloc(c->npad));
    PMQ_HIP(c->pgo.alloc(16 * (size_t)c->n));
    PMQ_HIP(c->uind.alloc(6 * (size_t)c->npad));
    const void *after[5] = {c->pg_dm.p, c->pg_dpd.p, c->pg_dth.p, c->pgo.p, c->uind.p};
    *moved = false;
    for (int k = 0; k < 5; k++) *moved = *moved || before[k] != after[k];
    return PMQ_OK;
}

static int batch_compute(pmq_ctx *c, int mode, unsigned terms, int precision)
{
    int rc = batch_check(c, mode, terms);
    if (rc) return rc;
    PMQ_HIP(hipSetDevice(c->device));
    c->last_mode = mode;
    c->last_terms = terms;
    c->last_precision = precision;
    c->pending = false;
    c->pcg_iters = 0;
    const bool pg = c->b_pgrad && mode == PMQ_GRADIENT;
    if (pg) {
        bool moved = false;
        if ((rc = batch_pgrad_buffers(c, &moved))) return rc;
        if (moved) batch_graph_release(c);
    }
    const int p = precision == PMQ_MIXED ? 1 : 0;
    if (!c->b_tuned[p] || c->b_tuned_pg[p] != pg) {
        // the autotuner: the shapes of the size buckets timed on this batch
        if ((rc = batch_tune(c, mode, terms & PMQ_TERM_MPOLE, terms & PMQ_TERM_POLAR, p == 1, pg))) return rc;
        c->b_tuned[p] = true;
        c->b_tuned_pg[p] = pg;
    }
    const int *var = c->b_var[p];
    const int key = (((((mode * 8 + (int)terms) * 2 + precision) * 2 + (pg ? 1 : 0)) * 2 + (c->b_merged[p] ? 1 : 0)) *
                         16 + var[0]) * 256 + var[1] * 16 + var[2];
    if (c->profiling) { // events are not captured
        rc = batch_enqueue(c, mode, terms, precision, var);
        c->b_pg_valid = pg && !rc;
        return rc;
    }
    if (c->b_graph_key != key) {
        batch_graph_release(c);
        PMQ_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = batch_enqueue(c, mode, terms, precision, var);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(c->stream, &g);
        if (rc) return rc;
        PMQ_HIP(ce);
        c->b_graph = g;
        hipGraphExec_t ex = nullptr;
        PMQ_HIP(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
        c->b_exec = ex;
        c->b_graph_key = key;
    }
    PMQ_HIP(hipGraphLaunch((hipGraphExec_t)c->b_exec, c->stream));
    c->b_pg_valid = pg;
    return PMQ_OK;
}

int pmq_set_batch_param_grad(pmq_ctx *c, int enable)
{
    CHECK_CTX(c);
    SETTLE_PENDING(c);
    new_epoch(c);
    batch_graph_release(c);
    c->b_pgrad = enable != 0;
    c->b_pg_valid = false;
    return PMQ_OK;
}

int pmq_set_batch_shape(pmq_ctx *c, int precision, int bucket, int lanes, int wgs_per_cu)
{
    CHECK_CTX(c);
    SETTLE_PENDING(c);
    if ((precision != PMQ_PLAIN_FP64 && precision != PMQ_MIXED) || bucket < 0 || bucket >= PMQ_BATCH_BUCKETS) {
        c->err = "pmq_set_batch_shape: precision PMQ_PLAIN_FP64 or PMQ_MIXED, bucket 0..2";
        return PMQ_ERR_ARG;
    }
    const int p = precision == PMQ_MIXED ? 1 : 0;
    if (lanes == 0) {
        c->b_pin[p][bucket] = false;
        c->b_tuned[p] = false; // the tuner decides again
        batch_graph_release(c);
        return PMQ_OK;
    }
    for (int v = 0; v < batch_variants(bucket); v++) {
        int r, l, w;
        batch_shape(bucket, v, &r, &l, &w);
        if (l == lanes && w == wgs_per_cu) {
            new_epoch(c);
            batch_graph_release(c);
            c->b_var[p][bucket] = v;
            c->b_ms[p][bucket] = -1.0f;
            c->b_pin[p][bucket] = true;
            c->b_merged[p] = false;
            return PMQ_OK;
        }
    }
    c->err = "pmq_set_batch_shape: no such shape for this bucket";
    return PMQ_ERR_ARG;
}

int pmq_batch_tuning(pmq_ctx *c, int precision, int *rows, int *lanes, int *wgs_per_cu, int *nclusters, float *ms)
{
    CHECK_CTX(c);
    if (precision != PMQ_PLAIN_FP64 && precision != PMQ_MIXED) {
        c->err = "pmq_batch_tuning: precision is PMQ_PLAIN_FP64 or PMQ_MIXED";
        return PMQ_ERR_ARG;
    }
    const int p = precision == PMQ_MIXED ? 1 : 0;
    const bool merged = c->b_tuned[p] && c->b_merged[p];
    for (int b = 0; b < PMQ_BATCH_BUCKETS; b++) {
        const int nb = c->nsys > 0 ? (merged ? (b == 0 ? c->nsys : 0) : c->b_beg[b + 1] - c->b_beg[b]) : 0;
        int r = -1, l = -1, w = -1;
        if (c->nsys > 0 && (c->b_tuned[p] || c->b_pin[p][b]) && nb > 0) batch_shape(b, c->b_var[p][b], &r, &l, &w);
        if (rows) rows[b] = c->nsys > 0 ? batch_bucket_rows(b) : -1;
        if (lanes) lanes[b] = l;
        if (wgs_per_cu) wgs_per_cu[b] = w;
        if (nclusters) nclusters[b] = nb;
        if (ms) ms[b] = (c->nsys > 0 && c->b_tuned[p] && nb > 0) ? c->b_ms[p][b] : -1.0f;
    }
    return PMQ_OK;
}

int pmq_set_batch(pmq_ctx *c, int nsys, const int *off)
{
    CHECK_CTX(c);
    SETTLE_PENDING(c);
    new_epoch(c);
    if (!c->have_top) {
        c->err = "pmq_set_batch: call pmq_set_topology first";
        return PMQ_ERR_STATE;
    }
    batch_graph_release(c);
    c->b_tuned[0] = c->b_tuned[1] = false;
    c->b_pg_valid = false;
    if (nsys <= 0) {
        c->nsys = 0;
        return PMQ_OK;
    }
    if (!off || off[0] != 0 || off[nsys] != c->n) {
        c->err = "pmq_set_batch: offsets must start at 0 and end at n";
        return PMQ_ERR_ARG;
    }
    std::vector<int> sizes(nsys), owner(c->n);
    for (int k = 0; k < nsys; k++) {
        sizes[k] = off[k + 1] - off[k];
        if (sizes[k] <= 0 || sizes[k] > 256) {
            c->err = "pmq_set_batch: every cluster needs 1..256 atoms";
            return PMQ_ERR_ARG;
        }
        for (int i = off[k]; i < off[k + 1]; i++) owner[i] = k;
    }
    // excluded pairs (sorted by first atom) must stay inside their cluster
    std::vector<int> eoff(nsys + 1, 0);
    for (const ExclPair &e : c->h_excl) {
        if (owner[e.i] != owner[e.k]) {
            c->err = "pmq_set_batch: a scaled pair joins two clusters";
            return PMQ_ERR_ARG;
        }
        eoff[owner[e.i] + 1]++;
    }
    for (int k = 0; k < nsys; k++) eoff[k + 1] += eoff[k];
    c->nsys = 0; // one system until the batch is fully uploaded
    std::vector<int> order(nsys);
    for (int k = 0; k < nsys; k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sizes[a] > sizes[b]; });
    // size buckets over the largest-first order: more than 128 atoms, more
    // than 64, the rest (each bucket's workgroups hold 256, 128, 64 rows)
    int beg[PMQ_BATCH_BUCKETS + 1] = {0, 0, 0, nsys};
    for (int k = 0; k < nsys; k++) {
        if (sizes[order[k]] > 128) beg[1] = k + 1;
        if (sizes[order[k]] > 64) beg[2] = k + 1;
    }
    int rc;
    if ((rc = upload(c, c->b_order, order.data(), nsys))) return rc;
    if ((rc = upload(c, c->b_off, off, nsys + 1))) return rc;
    if ((rc = upload(c, c->b_eoff, eoff.data(), nsys + 1))) return rc;
    PMQ_HIP(c->b_em.alloc(nsys)); PMQ_HIP(c->b_ep.alloc(nsys)); PMQ_HIP(c->b_it.alloc(nsys));
    PMQ_HIP(c->b_counter.alloc(PMQ_BATCH_BUCKETS));
    // streams and events of the side-by-side bucket launches (default priority)
    for (int k = 0; k < PMQ_BATCH_BUCKETS - 1; k++) {
        if (!c->b_side[k]) PMQ_HIP(hipStreamCreateWithFlags(&c->b_side[k], hipStreamNonBlocking));
        if (!c->b_join[k]) PMQ_HIP(hipEventCreateWithFlags(&c->b_join[k], hipEventDisableTiming));
    }
    if (!c->b_fork) PMQ_HIP(hipEventCreateWithFlags(&c->b_fork, hipEventDisableTiming));
    // frames run in caller order
    std::vector<int> iota(c->npad);
    for (int s = 0; s < c->npad; s++) iota[s] = s < c->n ? s : -1;
    if ((rc = upload(c, c->perm, iota.data(), c->npad))) return rc;
    PMQ_HIP(c->dep.alloc(3 * (size_t)c->n));
    PMQ_HIP(c->tep.alloc(3 * (size_t)c->n));
    PMQ_HIP(hipStreamSynchronize(c->stream));
    c->h_boff.assign(off, off + nsys + 1);
    for (int b = 0; b <= PMQ_BATCH_BUCKETS; b++) c->b_beg[b] = beg[b];
    c->nsys = nsys;
    return PMQ_OK;
}

int pmq_fetch_batch(pmq_ctx *c, double *em, double *ep, int *iters)
{
    CHECK_CTX(c);
    if (c->nsys <= 0) {
        c->err = "pmq_fetch_batch: no batch set";
        return PMQ_ERR_STATE;
    }
    if (c->noalloc_last) {
        c->err = "pmq_fetch_batch: the last evaluation (pmq_compute_batch_noalloc) wrote its results to the caller's "
                 "arrays";
        return PMQ_ERR_STATE;
    }
    PMQ_HIP(hipStreamSynchronize(c->stream));
    const int ns = c->nsys;
    if (em) PMQ_HIP(hipMemcpy(em, c->b_em.p, ns * sizeof(double), hipMemcpyDeviceToHost));
    if (ep) PMQ_HIP(hipMemcpy(ep, c->b_ep.p, ns * sizeof(double), hipMemcpyDeviceToHost));
    if (iters) {
        std::vector<double> it(ns);
        PMQ_HIP(hipMemcpy(it.data(), c->b_it.p, ns * sizeof(double), hipMemcpyDeviceToHost));
        for (int k = 0; k < ns; k++) iters[k] = (int)it[k];
    }
    return PMQ_OK;
}

int pmq_compute_batch_noalloc(pmq_ctx *c, int mode, unsigned terms, int precision, double *em, double *ep,
                              int *pcg_iters, double *etot, double *grad)
{
    CHECK_CTX(c);
    SETTLE_PENDING(c);
    new_epoch(c);
    int rc = check_compute(c, mode, terms, precision);
    if (rc) return rc;
    if (c->nsys <= 0) {
        c->err = "pmq_compute_batch_noalloc: no batch set (pmq_set_batch)";
        return PMQ_ERR_STATE;
    }
    if (c->nsys > PMQ_NOALLOC_MAX_CLUSTERS) {
        c->err = "pmq_compute_batch_noalloc: " + std::to_string(c->nsys) + " clusters exceed the fixed scratch (" +
                 std::to_string(PMQ_NOALLOC_MAX_CLUSTERS) + " clusters)";
        return PMQ_ERR_ARG;
    }
    if (c->profiling) {
        c->err = "pmq_compute_batch_noalloc: kernel timing creates events; switch profiling off first";
        return PMQ_ERR_STATE;
    }
    if (!c->static_scratch) {
        c->err = "pmq_compute_batch_noalloc: no static scratch address for this device";
        return PMQ_ERR_STATE;
    }
    if ((rc = batch_check(c, mode, terms))) return rc;
    PMQ_HIP(hipSetDevice(c->device));
    c->last_mode = mode;
    c->last_terms = terms;
    c->last_precision = precision;
    c->pcg_iters = 0;
    c->noalloc_last = true;
    c->b_pg_valid = false;
    // the shapes a previous pmq_compute tuned for this precision, else the first of each bucket
    const int p = precision == PMQ_MIXED ? 1 : 0;
    static const int first[PMQ_BATCH_BUCKETS] = {0, 0, 0};
    return batch_noalloc_enqueue(c, mode, terms & PMQ_TERM_MPOLE, terms & PMQ_TERM_POLAR, precision == PMQ_MIXED,
                                 c->b_tuned[p] ? c->b_var[p] : first, c->b_tuned[p] && c->b_merged[p], em, ep,
                                 pcg_iters, etot, grad);
}

// A speculative evaluation (no host round trips) of one system.  The first
// one with a given key is enqueued launch by launch; the second is captured
// into a HIP graph (every buffer it touches was sized and every host-side
// choice made by the first), and later ones with the same key replay it.
// Replays reproduce the host-side effects of compute_enqueue that the check
// reads.  A capture that fails leaves the key to plain enqueueing.
static int compute_speculative(pmq_ctx *c, int mode, unsigned terms, int precision)
{
    const pmq_ctx::EvalKey k = eval_key(c, mode, terms, precision);
    if (c->e_exec && same_key(k, c->e_key)) {
        PMQ_HIP(hipSetDevice(c->device));
        PMQ_HIP(hipGraphLaunch((hipGraphExec_t)c->e_exec, c->stream));
        c->pcg_iters = 0;
        c->last_mode = mode;
        c->last_terms = terms;
        c->last_precision = precision;
        return PMQ_OK;
    }
    if (c->e_last_ok && same_key(k, c->e_last) && !c->e_nocapture) {
        eval_graph_release(c);
        PMQ_HIP(hipSetDevice(c->device));
        PMQ_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = compute_enqueue(c, mode, terms, precision, false);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(c->stream, &g);
        hipGraphExec_t ex = nullptr;
        if (rc == PMQ_OK && ce == hipSuccess && g && hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess) {
            c->e_graph = g;
            c->e_exec = ex;
            c->e_key = k;
            PMQ_HIP(hipGraphLaunch(ex, c->stream));
            return PMQ_OK;
        }
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        c->e_nocapture = true;
    }
    const int rc = compute_enqueue(c, mode, terms, precision, false);
    if (rc) return rc;
    c->e_last = k;
    c->e_last_ok = true;
    return PMQ_OK;
}

int pmq_compute(pmq_ctx *c, int mode, unsigned terms, int precision)
{
    CHECK_CTX(c);
    int rc = check_compute(c, mode, terms, precision);
    if (rc) return rc;
    c->noalloc_last = false;
    if (c->nsys > 0) return batch_compute(c, mode, terms, precision);
    c->pending = false;
    if (c->safe || !c->graphs_on || c->profiling) {
        if (c->safe) new_epoch(c); // a safe run sizes buffers
        rc = compute_enqueue(c, mode, terms, precision, c->safe);
    } else {
        rc = compute_speculative(c, mode, terms, precision);
    }
    if (rc) return rc;
    c->pending = true;
    c->pending_safe = c->safe;
    return PMQ_OK;
}

int pmq_set_graphs(pmq_ctx *c, int enable)
{
    CHECK_CTX(c);
    SETTLE_PENDING(c);
    new_epoch(c);
    eval_graph_release(c);
    c->graphs_on = enable != 0;
    return PMQ_OK;
}

// wait for the evaluation and check what was decided without the device:
// buffer capacities and the speculative PCG batch.  A run that fell short
// is redone with host round trips (and the capacities it measured).
static int pmq_check(pmq_ctx *c)
{
    PMQ_HIP(hipSetDevice(c->device));
    PMQ_HIP(hipStreamSynchronize(c->stream));
    if (!c->pending) return PMQ_OK;
    c->pending = false;
    const HostStatus h = *c->h_status;
    const bool pol = c->last_terms & PMQ_TERM_POLAR;
    const bool over = h.st.overflow != 0;
    const bool short_pcg = pol && !c->pending_safe && h.sc[8] == 0.0;
    if (!over && !short_pcg) {
        if (pol && !c->pending_safe) {
            c->pcg_iters = (int)h.sc[9];
            c->pcg_guess_iters = c->pcg_iters;
        }
        c->safe = false;
        return PMQ_OK;
    }
    if (h.st.ntiles > c->tile_cap) c->tile_cap = 0;      // resized from the count on the safe run
    if ((size_t)h.st.nnz > c->tcap) c->tcap = 0;
    new_epoch(c);
    int rc = compute_enqueue(c, c->last_mode, c->last_terms, c->last_precision, true);
    if (rc) return rc;
    PMQ_HIP(hipStreamSynchronize(c->stream));
    c->safe = false;
    return PMQ_OK;
}

int pmq_param_grad(pmq_ctx *c, double *dpole, double *dpolarity, double *dpdamp, double *dthole)
{
    CHECK_CTX(c);
    int rc = pmq_check(c);
    if (rc) return rc;
    if (!c->have_top || !c->have_prm || !c->have_xyz) {
        c->err = "pmq_param_grad: topology, parameters and coordinates must be set and evaluated first";
        return PMQ_ERR_STATE;
    }
    if (c->nsys > 0) {
        // batched clusters: computed inside the evaluation's own launch
        if (!c->b_pg_valid || c->noalloc_last) {
            c->err = "pmq_param_grad: a batch needs pmq_set_batch_param_grad(ctx, 1) and then a GRADIENT evaluation";
            return PMQ_ERR_STATE;
        }
        const size_t n = c->n;
        const bool pol = (c->last_terms & PMQ_TERM_POLAR) && !c->h_polarity.empty();
        double *outs[4] = {dpole, dpolarity, dpdamp, dthole};
        const size_t at[4] = {0, 13 * n, 14 * n, 15 * n}, len[4] = {13 * n, n, n, n};
        for (int k = 0; k < 4; k++) {
            if (!outs[k]) continue;
            if (k > 0 && !pol) {
                for (size_t i = 0; i < n; i++) outs[k][i] = 0.0;
                continue;
            }
            PMQ_HIP(hipMemcpy(outs[k], c->pgo.p + at[k], len[k] * sizeof(double), hipMemcpyDeviceToHost));
        }
        return PMQ_OK;
    }
    if (c->last_mode != PMQ_GRADIENT || c->last_precision != PMQ_PLAIN_FP64) {
        c->err = "pmq_param_grad: needs a plain-FP64 GRADIENT evaluation of one system first";
        return PMQ_ERR_STATE;
    }
    if (c->prm.pentype != PMQ_PEN_NONE || c->prm.use_chgflx) {
        c->err = "pmq_param_grad: charge penetration and charge flux are not differentiated";
        return PMQ_ERR_UNSUPPORTED;
    }
    new_epoch(c); // its buffers may grow
    return param_gradients(c, c->last_terms & PMQ_TERM_MPOLE, c->last_terms & PMQ_TERM_POLAR, dpole, dpolarity,
                           dpdamp, dthole);
}

int pmq_sync(pmq_ctx *c)
{
    CHECK_CTX(c);
    return pmq_check(c);
}

int pmq_fetch(pmq_ctx *c, pmq_result *r)
{
    CHECK_CTX(c);
    if (!r) return PMQ_ERR_ARG;
    int rc0 = pmq_check(c);
    if (rc0) return rc0;
    if (c->noalloc_last) {
        c->err = "pmq_fetch: the last evaluation (pmq_compute_batch_noalloc) wrote its results to the caller's arrays";
        return PMQ_ERR_STATE;
    }
    Totals t;
    PMQ_HIP(hipMemcpy(&t, c->totals.p, sizeof t, hipMemcpyDeviceToHost));
    const int n = c->n;
    const bool grad = c->last_mode == PMQ_GRADIENT || c->last_mode == PMQ_VIRIAL;
    r->em = t.em;
    r->ep = t.ep;
    r->er = t.er;
    r->edsp = t.edsp;
    r->ev = t.ev;
    r->ect = t.ect;
    r->esum = t.em + t.ep + t.er + t.edsp + t.ev + t.ect;
    // with Ewald the kernels sum the intramolecular pairs' own energies:
    // the intermolecular part is the rest of em
    const bool ewald_ana = c->last_mode == PMQ_ANALYSIS && c->prm.use_ewald && (c->last_terms & PMQ_TERM_MPOLE);
    r->einter = ewald_ana ? t.em - t.einter : t.einter;
    r->nem = (long long)t.nem;
    r->nep = c->last_mode == PMQ_ANALYSIS ? (long long)t.nep : 0;
    for (int k = 0; k < 9; k++) r->vir[k] = c->last_mode == PMQ_VIRIAL ? t.vir[k] : 0.0;
    r->pcg_iters = c->pcg_iters;
    const bool mp = c->last_terms & PMQ_TERM_MPOLE, pol = c->last_terms & PMQ_TERM_POLAR;
    std::vector<double> gm, gp;
    if (grad && mp) {
        gm.resize(3 * (size_t)n);
        PMQ_HIP(hipMemcpy(gm.data(), c->dem.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
    }
    if (grad && pol) {
        gp.resize(3 * (size_t)n);
        PMQ_HIP(hipMemcpy(gp.data(), c->dep.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
    }
    if (r->dem && grad) {
        for (size_t j = 0; j < 3 * (size_t)n; j++) r->dem[j] = mp ? gm[j] : 0.0;
    }
    if (r->dep && grad) {
        for (size_t j = 0; j < 3 * (size_t)n; j++) r->dep[j] = pol ? gp[j] : 0.0;
    }
    if (r->desum && grad) {
        for (size_t j = 0; j < 3 * (size_t)n; j++) r->desum[j] = (mp ? gm[j] : 0.0) + (pol ? gp[j] : 0.0);
    }
    const bool rp = c->last_terms & PMQ_TERM_REPEL, dsp = c->last_terms & PMQ_TERM_DISP;
    std::vector<double> gr, gd;
    if (grad && rp) {
        gr.resize(3 * (size_t)n);
        PMQ_HIP(hipMemcpy(gr.data(), c->der.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
    }
    if (grad && dsp) {
        gd.resize(3 * (size_t)n);
        PMQ_HIP(hipMemcpy(gd.data(), c->dedsp.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
    }
    if (r->der && grad)
        for (size_t j = 0; j < 3 * (size_t)n; j++) r->der[j] = rp ? gr[j] : 0.0;
    if (r->dedsp && grad)
        for (size_t j = 0; j < 3 * (size_t)n; j++) r->dedsp[j] = dsp ? gd[j] : 0.0;
    if (r->desum && grad)
        for (size_t j = 0; j < 3 * (size_t)n; j++) r->desum[j] += (rp ? gr[j] : 0.0) + (dsp ? gd[j] : 0.0);
    const bool vd = c->last_terms & PMQ_TERM_VDW, ct = c->last_terms & PMQ_TERM_CTRN;
    std::vector<double> gv, gc;
    if (grad && vd) {
        gv.resize(3 * (size_t)n);
        PMQ_HIP(hipMemcpy(gv.data(), c->dev.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
    }
    if (grad && ct) {
        gc.resize(3 * (size_t)n);
        PMQ_HIP(hipMemcpy(gc.data(), c->dect.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
    }
    if (r->dev && grad)
        for (size_t j = 0; j < 3 * (size_t)n; j++) r->dev[j] = vd ? gv[j] : 0.0;
    if (r->dect && grad)
        for (size_t j = 0; j < 3 * (size_t)n; j++) r->dect[j] = ct ? gc[j] : 0.0;
    if (r->desum && grad)
        for (size_t j = 0; j < 3 * (size_t)n; j++) r->desum[j] += (vd ? gv[j] : 0.0) + (ct ? gc[j] : 0.0);
    if (pol && r->uind) PMQ_HIP(hipMemcpy(r->uind, c->uind_out.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
    if (pol && r->uinp)
        PMQ_HIP(hipMemcpy(r->uinp, c->uind_out.p + 3 * (size_t)n, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
    if (r->aem && c->last_mode == PMQ_ANALYSIS)
        PMQ_HIP(hipMemcpy(r->aem, c->aem.p, n * sizeof(double), hipMemcpyDeviceToHost));
    if (r->aep && c->last_mode == PMQ_ANALYSIS && pol)
        PMQ_HIP(hipMemcpy(r->aep, c->aep.p, n * sizeof(double), hipMemcpyDeviceToHost));
    if (r->rpole) PMQ_HIP(hipMemcpy(r->rpole, c->rpole13.p, 13 * n * sizeof(double), hipMemcpyDeviceToHost));
    if (!std::isfinite(r->esum)) {
        c->err = "pmq_fetch: non-finite total energy";
        return PMQ_ERR_NAN;
    }
    if (pol && c->pcg_iters < 0) {
        c->err = "pmq_fetch: induced dipoles did not converge within politer iterations";
        return PMQ_ERR_NOCONV;
    }
    if (pol && c->nsys > 0) {
        // batched clusters: each cluster's iteration count, negative when
        // its solve ran out of politer iterations
        std::vector<double> it(c->nsys);
        PMQ_HIP(hipMemcpy(it.data(), c->b_it.p, c->nsys * sizeof(double), hipMemcpyDeviceToHost));
        for (int k = 0; k < c->nsys; k++)
            if (it[k] < 0.0) {
                c->err = "pmq_fetch: the induced dipoles of cluster " + std::to_string(k) +
                         " did not converge within politer iterations";
                return PMQ_ERR_NOCONV;
            }
    }
    return PMQ_OK;
}

int pmq_eval(pmq_ctx *c, int mode, unsigned terms, int precision, const double *x, const double *y, const double *z,
             const pmq_box *box, pmq_result *res)
{
    int rc = pmq_set_coords(c, x, y, z, box);
    if (rc) return rc;
    if ((rc = pmq_compute(c, mode, terms, precision))) return rc;
    return pmq_fetch(c, res);
}

} // extern "C"
