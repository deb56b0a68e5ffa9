// Host runtime of the gfx950 NMF library: contexts, device-resident
// matrices and solves, the iteration driver (captured into a HIP graph),
// and the C ABI declared in include/nmfgpu.h.
//
// Iteration driver: rust/src/lib.rs nmf_run_lee (1588-1647),
// nmf_run_brunet (1528-1586), nmf_run_hals (637-688), nmf_run_ehals
// (710-790), nmf_run_dnmf (839-906), nmf_run_offset (1649-1759),
// nmf_run_nsnmf (1761-1833); the stationary stop follows
// StationaryStop (lib.rs:1336-1401), evaluated on the device (k_objfin).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <map>
#include <chrono>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "nmfg_fcnnls.h"
#include "nmfg_internal.h"
#include "nmfgpu.h"

using namespace nmfg;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess)                                                          \
            return fail(NMFG_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

int env_int(const char *name, int dflt) {
    const char *v = std::getenv(name);
    return v ? std::atoi(v) : dflt;
}

bool valid_form(int32_t form) {
    return form == NMFG_FP64 || form == NMFG_FP32_STORE || form == NMFG_BF16X2_STORE ||
           form == NMFG_FP32 || form == NMFG_BF16X2;
}

// Storage of V (VP_*) of a precision form.
int storage_of(int32_t form) {
    return form == NMFG_FP64 ? VP_F64
                             : (form == NMFG_BF16X2_STORE || form == NMFG_BF16X2 ? VP_B16X2 : VP_F32);
}

// Bytes of V's device block: V in its storage form, and for the BF16 form
// the column pass's copy (two planes of b16_rows(n) x ldv bfloat16) after
// it, 256-byte aligned.
size_t v_block_bytes(int form, int64_t n, int64_t ldv, size_t *vc_offset) {
    const size_t vb = (size_t)v_bytes_per_elem(storage_of(form)) * (size_t)n * ldv;
    const size_t off = (vb + 255) / 256 * 256;
    if (vc_offset) *vc_offset = off;
    return form == NMFG_BF16X2 ? off + 4 * (size_t)b16_rows(n) * ldv : vb;
}

int kind_of(int32_t algo) {
    switch (algo) {
        case NMFG_LEE: return K_LEE;
        case NMFG_BRUNET: return K_BRUNET;
        case NMFG_HALS: return K_HALS;
        case NMFG_DNMF: return K_DNMF;
        case NMFG_EHALS: return K_EHALS;
        case NMFG_OFFSET: return K_OFFSET;
        case NMFG_NSNMF: return K_NSNMF;
        case NMFG_LSNMF: return K_LSNMF;
        default: return -1;
    }
}

}  // namespace

namespace {
// Per device: the address of the no-allocation entry point's scratch array
// (resolved once, when the first context on the device is created) and the
// lock that serialises that entry point's calls on the device.
constexpr int kMaxDevices = 64;
struct NoallocDevice {
    std::once_flag once;
    hipError_t err = hipErrorNotInitialized;
    void *ptr = nullptr;
    std::mutex mu;
};
NoallocDevice g_noalloc[kMaxDevices];
}  // namespace

// A device-memory budget shared by contexts (serving workers): every block
// a context holds, in use or idle in its cache, is charged to it.
// Lock order: list_mu, then a context's pool, then mu.
struct nmfg_budget {
    std::mutex list_mu;            // guards ctxs
    std::vector<nmfg_ctx *> ctxs;  // attached contexts
    std::mutex mu;                 // guards the three counts
    int64_t cap = 0;
    int64_t used = 0;
    int64_t peak = 0;
};

struct nmfg_ctx {
    int device = 0;
    nmfg_budget *budget = nullptr;
    hipStream_t stream = nullptr;
    std::recursive_mutex mu;
    // Device blocks of the context (ctx_alloc / ctx_release): released
    // blocks wait in `idle`, by size, for the next allocation.  `pool`
    // guards the three containers below, so that another context's thread
    // may free idle blocks for a shared budget while this one is busy.
    std::mutex pool;
    std::multimap<size_t, void *> idle;
    std::unordered_map<void *, size_t> block_bytes;  // every block the context holds
    size_t idle_bytes = 0;
    size_t idle_cap = 0;
    bool poison = false;  // fill every block handed out with 0xff bytes (NaN)
};

struct nmfg_matrix {
    nmfg_ctx *ctx = nullptr;
    int64_t n = 0, p = 0, ldv = 0;
    int form = NMFG_FP64;       // precision form (enum nmfg_precision)
    int precision = VP_F64;     // storage of V (VP_*)
    void *V = nullptr;
    int64_t vplane = 0;         // VP_B16X2: elements from the hi to the lo plane
    uint16_t *Vc = nullptr;     // BF16 form: the column pass's copy (PassArgs::Vc), in V's block
    int64_t nc = 0;             // BF16 form: rows of Vc (b16_rows(n))
    nmfg_solve *prod[33] = {};  // cached K_PROD solves by chunk width
    double *mul_buf = nullptr;  // products: B then C, row-major, on the device
    size_t mul_cap = 0;         // doubles
};

// A captured chunk of iterations (instantiated graph) and whether capture
// failed for it.
struct ChunkGraph {
    hipGraphExec_t exec = nullptr;
    bool failed = false;
};

struct nmfg_solve {
    nmfg_matrix *m = nullptr;
    nmfg_params prm{};
    int r = 0, R = 0, kind = 0;
    int batch = 1;  // independent solves sharing V (restarts), see Batch
    const RankKernels *K = nullptr;
    double *W0 = nullptr, *H0 = nullptr;
    double *W = nullptr, *Wx = nullptr, *H = nullptr, *Hp = nullptr;
    double *red = nullptr, *hst = nullptr, *wsx = nullptr, *wpart = nullptr, *wsc = nullptr;
    double *obj_part = nullptr;
    double *gpart = nullptr, *gram = nullptr, *rpart = nullptr;
    double *off0 = nullptr, *off = nullptr;  // offset: initial and working offset (n per solve)
    double *rsv = nullptr;                   // offset: row sums of V (n)
    void *wgt = nullptr;                     // lsNMF: weight matrix (n x ldv, form wf)
    int wf = WF_F64;
    bool xgram = false;    // column pass takes the W Gram from k_wgram
    bool rowmfma = false;  // MFMA row pass (column splits) + rowfin
    int nsplit = 1, nrb = 0;
    int row_tiles = 1;     // MFMA row pass: 16-row tiles per block
    int row_waves = NWR;   // VALU row pass waves per work-group
    int kl_split = 1;      // KL column pass row splits
    double *kpart = nullptr;
    // objective-based stops, decided on the device
    int *d_done = nullptr;
    int64_t *d_niter = nullptr;  // [stop iteration, evaluation counter]
    hipGraphExec_t chunk_graph = nullptr;  // STOP_CHECK_CHUNK iterations + objectives
    bool chunk_graph_failed = false;
    // brunet with an objective-based stop: the KL column pass also sums the
    // previous iteration's objective (fuse_obj); Hs keeps H(t) until the stop
    // test of iteration t has passed
    bool fuse_obj = false, fused_active = false;
    double *Hs = nullptr;
    double *d_devs = nullptr, *d_state = nullptr;
    const int *loop_done = nullptr;  // passed to the loop's kernels while a stop is armed
    int ntiles_col = 0, ntiles_h = 0, ntiles_obj = 0, nrowblk = 0, nparts = 0, ngparts = 0;
    int osplit = 1;  // objective row splits per tile
    bool obj_mfma = false;  // objective pass with W H on the FP64 matrix cores
    bool have_factors = false;
    bool no_graph = false;  // enqueue kernels directly, never capture (no-allocation entry)
    // FP32 form: products on FP32 MFMA with the float copies of the factors
    bool f32 = false;
    // BF16 form (lee): products on bf16 MFMA over hi + lo copies of V
    // (m->Vc, V's planes), W (Wb, transposed) and H' (Hb); numerators per
    // column-pass row split (kl_split) in num
    bool b16 = false;
    uint16_t *Wb = nullptr, *Hb = nullptr;
    float *Wf = nullptr, *Hf = nullptr, *num = nullptr;
    // ragged batch (rank sweep): each member's rank (r is their maximum),
    // on the host and the device
    std::vector<int> mranks;
    int *d_ranks = nullptr;
    hipGraphExec_t graph = nullptr;
    int64_t n_iter = 0;
    std::vector<double> devs;
    std::vector<void *> allocs;
};

namespace {

struct Guard {
    nmfg_ctx *c;
    std::lock_guard<std::recursive_mutex> lk;
    explicit Guard(nmfg_ctx *ctx) : c(ctx), lk(ctx->mu) { (void)hipSetDevice(ctx->device); }
};

// Device memory of a context.  Every kernel and copy that touches the
// memory of a context's matrices and solves runs on the context's stream,
// so a released block can serve the context's next allocation at once: the
// work queued after that allocation runs after the work queued before the
// release.  hipFree waits for every stream of the device; with one context
// per serving thread, freeing each buffer would hold one thread's clean-up
// until the other threads' queued iterations end.  Released blocks stay
// idle, up to idle_cap bytes per context (NMFG_CACHE_MB, default 8192 MiB;
// 0 frees at once), until the context is destroyed or
// nmfg_ctx_release_cached is called.
size_t align256(size_t b) { return (b + 255) / 256 * 256; }

size_t block_class(size_t bytes) {
    // four sizes per power of two: a block is at most 25 % above the request
    const size_t b = std::max<size_t>(bytes, 256);
    const size_t step = (size_t)1 << (61 - __builtin_clzll(b));
    return (b + step - 1) / step * step;
}

void budget_charge(nmfg_budget *bg, int64_t bytes) {
    if (!bg) return;
    std::lock_guard<std::mutex> lk(bg->mu);
    bg->used += bytes;
    bg->peak = std::max(bg->peak, bg->used);
}

// Charges `bytes` to the budget if they fit under its cap.
bool budget_reserve(nmfg_budget *bg, size_t bytes) {
    std::lock_guard<std::mutex> lk(bg->mu);
    if (bg->used + (int64_t)bytes > bg->cap) return false;
    bg->used += (int64_t)bytes;
    bg->peak = std::max(bg->peak, bg->used);
    return true;
}

// Frees a block the context holds (uncharging its budget); pool held.
void ctx_free_block(nmfg_ctx *c, void *p, size_t b) {
    if (hipFree(p) != hipSuccess) (void)hipGetLastError();  // not left for the next launch check
    c->block_bytes.erase(p);
    budget_charge(c->budget, -(int64_t)b);
}

// Frees the idle blocks; pool held.  hipFree waits for the device, so work
// queued before a block was released has finished with it.
void ctx_trim_locked(nmfg_ctx *c) {
    for (auto &kv : c->idle) ctx_free_block(c, kv.second, kv.first);
    c->idle.clear();
    c->idle_bytes = 0;
}

void ctx_trim(nmfg_ctx *c) {
    std::lock_guard<std::mutex> lk(c->pool);
    ctx_trim_locked(c);
}

// Charges `bytes` to the context's budget, freeing idle blocks to make room:
// the context's own first, then those of the other contexts on the budget
// (idle blocks are unused whatever their context is doing).
bool budget_make_room(nmfg_ctx *c, size_t bytes) {
    nmfg_budget *bg = c->budget;
    if (!bg) return true;
    if (budget_reserve(bg, bytes)) return true;
    ctx_trim(c);
    if (budget_reserve(bg, bytes)) return true;
    std::lock_guard<std::mutex> ll(bg->list_mu);
    // Other threads keep releasing blocks meanwhile (a request's product
    // solves and staging rows go idle while it runs), and a block released
    // into a context this loop has already passed stays charged: pass over
    // every context again until a whole pass finds no idle block.  Then
    // every charged byte is in use by admitted work, and a request within
    // its admitted footprint fits.
    for (;;) {
        bool freed = false;
        for (nmfg_ctx *o : bg->ctxs) {
            {
                std::lock_guard<std::mutex> lk(o->pool);
                if (o->idle.empty()) continue;
                (void)hipSetDevice(o->device);
                ctx_trim_locked(o);
                (void)hipSetDevice(c->device);
                freed = true;
            }
            if (budget_reserve(bg, bytes)) return true;
        }
        if (budget_reserve(bg, bytes)) return true;
        if (!freed) return false;
    }
}

int ctx_alloc(nmfg_ctx *c, size_t bytes, void **out) {
    const size_t b = block_class(bytes);
    *out = nullptr;
    {
        std::lock_guard<std::mutex> lk(c->pool);
        auto it = c->idle.lower_bound(b);
        // under a budget only a block of the same class is reused, so the
        // bytes in use never exceed the sum of the requests' classes
        const size_t slack = c->budget ? 0 : b / 2;
        if (it != c->idle.end() && it->first <= b + slack) {
            *out = it->second;
            c->idle_bytes -= it->first;
            c->idle.erase(it);
        }
    }
    if (!*out) {
        if (!budget_make_room(c, b))
            return fail(NMFG_EBUDGET, "device memory budget exhausted: " + std::to_string(b) +
                                          " more bytes do not fit");
        hipError_t e = hipMalloc(out, b);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            std::lock_guard<std::mutex> lk(c->pool);
            if (!c->idle.empty()) {
                ctx_trim_locked(c);
                e = hipMalloc(out, b);
            }
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            *out = nullptr;
            budget_charge(c->budget, -(int64_t)b);
            return fail(NMFG_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        std::lock_guard<std::mutex> lk(c->pool);
        c->block_bytes[*out] = b;
    }
    if (c->poison) HIPCHK(hipMemsetAsync(*out, 0xff, b, c->stream));
    return NMFG_OK;
}

void ctx_release(nmfg_ctx *c, void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(c->pool);
    auto it = c->block_bytes.find(p);
    if (it == c->block_bytes.end()) return;
    const size_t b = it->second;
    if (b > c->idle_cap) {
        ctx_free_block(c, p, b);
        return;
    }
    c->idle.emplace(b, p);
    c->idle_bytes += b;
    while (c->idle_bytes > c->idle_cap) {  // over the limit: free the largest
        auto big = std::prev(c->idle.end());
        ctx_free_block(c, big->second, big->first);
        c->idle_bytes -= big->first;
        c->idle.erase(big);
    }
}

int dmalloc(nmfg_solve *s, void **p, size_t bytes) {
    int rc = ctx_alloc(s->m->ctx, bytes, p);
    if (rc) return rc;
    s->allocs.push_back(*p);
    return NMFG_OK;
}

void solve_free(nmfg_solve *s) {
    if (s->graph) (void)hipGraphExecDestroy(s->graph);
    if (s->chunk_graph) (void)hipGraphExecDestroy(s->chunk_graph);
    s->graph = nullptr;
    s->chunk_graph = nullptr;
    for (void *p : s->allocs) ctx_release(s->m->ctx, p);
    s->allocs.clear();
}

double solve_eps(const nmfg_solve *s) {
    return s->kind == K_BRUNET ? DBL_EPSILON : s->prm.eps;  // lib.rs:1539
}

// Per-solve buffer sizes (doubles) of a batched solve.
int nrec_of(const nmfg_solve *s) { return std::max(s->ntiles_col, s->ntiles_h); }

Batch batch_of(const nmfg_solve *s) {
    Batch b{};
    const int R = s->R;
    b.count = s->batch;
    b.w = s->m->n * R;
    b.h = (int64_t)R * s->m->ldv;
    b.red = (int64_t)nrec_of(s) * nred(R);
    b.hst = nred(R);
    b.wsx = R;
    b.wpart = (int64_t)s->nparts * 2 * R;
    b.wsc = 3 * R;
    b.gpart = (int64_t)s->ngparts * nred(R);
    b.gram = nred(R);
    const int acc = s->kind == K_LSNMF ? 2 * R : R;  // accumulated values per row / column
    b.rpart = (int64_t)s->nsplit * s->m->n * acc;
    b.off = s->m->n;
    // KL: then kl_split x R partial column sums of W (one per split)
    b.kpart = (int64_t)s->kl_split * s->ntiles_col * acc * TJK +
              (kl_kind(s->kind) ? (int64_t)s->kl_split * R : 0);
    b.num = (int64_t)(s->b16 ? s->kl_split : 1) * R * s->m->ldv;
    b.wb = 2 * (int64_t)R * s->m->nc;
    b.hb = 2 * (int64_t)R * s->m->ldv;
    b.ranks = s->mranks.empty() ? nullptr : s->d_ranks;
    return b;
}

PassArgs pass_args(nmfg_solve *s, int first, int clamp_w, int clamp_h) {
    PassArgs a{};
    nmfg_matrix *m = s->m;
    a.V = m->V;
    a.vplane = m->vplane;
    a.ldv = m->ldv;
    a.n = (int)m->n;
    a.p = (int)m->p;
    a.r = s->r;
    a.ntiles = s->ntiles_col;
    a.nrowblk = s->nrowblk;
    a.W = s->W;
    a.Wx = s->Wx;
    a.H = s->H;
    a.Hp = s->Hp;
    a.red = s->red;
    a.wsx = s->wsx;
    a.hst = s->hst;
    a.eps = solve_eps(s);
    a.rescale = s->prm.rescale ? 1 : 0;
    a.first = first;
    a.clamp_w = clamp_w;
    a.clamp_h = clamp_h;
    a.kind = s->kind;
    a.bt = batch_of(s);
    a.gram = s->gram;
    a.xgram = s->xgram ? 1 : 0;
    a.rpart = s->rpart;
    a.nsplit = s->nsplit;
    a.nrb = s->nrb;
    a.rowmfma = s->rowmfma ? 1 : 0;
    a.row_tiles = s->row_tiles;
    a.opart = s->fused_active ? s->obj_part : nullptr;
    a.done = s->loop_done;
    a.row_waves = s->row_waves;
    a.off = s->kind == K_OFFSET ? s->off : nullptr;
    a.rsv = s->rsv;
    a.kl_split = s->kl_split;
    a.kpart = s->kl_split > 1 && !s->b16 ? s->kpart : nullptr;
    a.wgt = s->wgt;
    a.wf = s->wf;
    a.theta = s->prm.theta;
    a.f32 = s->f32 ? 1 : 0;
    a.Wf = s->f32 ? s->Wf : nullptr;
    a.Hf = s->f32 ? s->Hf : nullptr;
    a.num = (s->f32 && !kl_kind(s->kind)) || s->b16 ? s->num : nullptr;
    a.b16 = s->b16 ? 1 : 0;
    a.Vc = s->b16 ? m->Vc : nullptr;
    a.cplane = s->b16 ? m->nc * m->ldv : 0;
    a.nc = (int)m->nc;
    a.Wb = s->b16 ? s->Wb : nullptr;
    a.Hb = s->b16 ? s->Hb : nullptr;
    return a;
}

// Gram record of the W the column pass streams (xgram solves).
int launch_wgram(nmfg_solve *s, hipStream_t st) {
    WGramArgs g{};
    g.W = s->kind == K_EHALS ? s->Wx : s->W;
    g.off = s->kind == K_OFFSET ? s->off : nullptr;
    g.n = (int)s->m->n;
    g.part = s->gpart;
    g.nparts = s->ngparts;
    g.out = s->gram;
    g.kind = s->kind;
    g.Wf = s->f32 ? s->Wf : nullptr;
    g.Wb = s->b16 ? s->Wb : nullptr;
    g.nc = (int)s->m->nc;
    g.bt = batch_of(s);
    g.done = s->loop_done;
    HIPCHK(s->K->wgram(g, st));
    return NMFG_OK;
}

int launch_wstats(nmfg_solve *s, hipStream_t st, int first, int dnmf_norm, bool parts_only = false);

// nsNMF: col_rescale of the stored W (divisors 1 on the first iteration) and
// Wx = W S, before the column pass streams Wx.
int launch_nsprep(nmfg_solve *s, hipStream_t st, int first) {
    int rc = launch_wstats(s, st, first, 0);
    if (rc) return rc;
    NsPrepArgs a{s->W, s->Wx, s->wsc, (int)s->m->n, s->r, s->prm.theta, batch_of(s), s->loop_done};
    HIPCHK(s->K->nsprep(a, st));
    return NMFG_OK;
}

int launch_col(nmfg_solve *s, hipStream_t st, int first, int clamp_h) {
    int rc = NMFG_OK;
    if (s->xgram && (rc = launch_wgram(s, st))) return rc;
    if (s->kind == K_NSNMF && (rc = launch_nsprep(s, st, first))) return rc;
    PassArgs a = pass_args(s, first, 0, clamp_h);
    a.lambda = s->prm.sparsity_h;
    a.gamma = s->prm.smoothness_h;
    if (s->kind == K_LSNMF) {
        HIPCHK(s->K->ls(a, st, 0));
        return NMFG_OK;
    }
    HIPCHK(s->K->col(a, st));
    // split KL passes: the partials' finisher (the BF16 form's numerator
    // splits are summed by its own H-update kernel, launched by col)
    if (s->kl_split > 1 && kl_kind(s->kind)) HIPCHK(s->K->colfin(a, st));
    return NMFG_OK;
}

int launch_red(nmfg_solve *s, hipStream_t st, int ntiles) {
    RedArgs a{s->red, ntiles, s->hst, batch_of(s), s->loop_done};
    HIPCHK(s->K->red(a, st));
    return NMFG_OK;
}

int launch_row(nmfg_solve *s, hipStream_t st, int clamp_w) {
    PassArgs a = pass_args(s, 0, clamp_w, 0);
    if (s->kind == K_NSNMF) a.H = s->Hp;  // the W update reads S H' (lib.rs:1813-1814)
    a.lambda = s->prm.sparsity_w;
    a.gamma = s->prm.smoothness_w;
    if (s->kind == K_LSNMF) {
        HIPCHK(s->K->ls(a, st, 1));
        return NMFG_OK;
    }
    HIPCHK(s->K->row(a, st));
    if (s->rowmfma || s->nsplit > 1) HIPCHK(s->K->rowfin(a, st));
    return NMFG_OK;
}

int launch_prep(nmfg_solve *s, hipStream_t st) {
    PrepArgs a{s->W, s->Wx, s->wsx, s->f32 ? s->Wf : nullptr, (int)s->m->n, s->r, s->kind,
               solve_eps(s), batch_of(s)};
    HIPCHK(s->K->prep(a, st));
    return NMFG_OK;
}

// H statistics of the stored H (records + reduction), for a W update that
// does not follow a column pass.
int launch_hstats(nmfg_solve *s, hipStream_t st) {
    if (s->batch != 1) return fail(NMFG_EUNSUPPORTED, "H statistics of a batched solve");
    HRecArgs a{s->H, s->m->ldv, (int)s->m->p, s->ntiles_h, s->red};
    HIPCHK(s->K->hrec(a, st));
    return launch_red(s, st, s->ntiles_h);
}

int launch_wstats(nmfg_solve *s, hipStream_t st, int first, int dnmf_norm, bool parts_only) {
    WStatArgs a{};
    a.W = s->W;
    a.n = (int)s->m->n;
    a.r = s->r;
    a.part = s->wpart;
    a.nparts = s->nparts;
    a.out = s->wsc;
    a.kind = s->kind;
    a.rescale = s->prm.rescale ? 1 : 0;
    a.first = first;
    a.dnmf_norm = dnmf_norm;
    a.parts_only = parts_only ? 1 : 0;
    a.bt = batch_of(s);
    a.done = s->loop_done;
    HIPCHK(s->K->wstats(a, st));
    return NMFG_OK;
}

int launch_fin(nmfg_solve *s, hipStream_t st, int clamp_h, int dnmf_fold) {
    Finalize f{};
    f.W = s->W;
    f.H = s->H;
    f.ldv = s->m->ldv;
    f.n = (int)s->m->n;
    f.p = (int)s->m->p;
    f.r = s->r;
    f.wsc = s->wsc;
    f.hst = s->hst;
    f.kind = s->kind;
    f.clamp_h = clamp_h;
    f.dnmf_fold = dnmf_fold;
    f.eps = solve_eps(s);
    f.bt = batch_of(s);
    HIPCHK(s->K->fin(f, st));
    return NMFG_OK;
}

// brunet: iteration `it` (1-based) clamps W' and H' when it % 10 == 0
// (lib.rs:1572-1575); H's clamp is applied when H is next read.
bool clamp_iter(const nmfg_solve *s, int64_t it) { return s->kind == K_BRUNET && it > 0 && it % 10 == 0; }

int enqueue_iteration(nmfg_solve *s, hipStream_t st, int64_t it) {
    int rc;
    if ((rc = launch_col(s, st, it == 1, clamp_iter(s, it - 1)))) return rc;
    // lsNMF's W update reads no statistics of H'
    if (s->kind != K_LSNMF && (rc = launch_red(s, st, s->ntiles_col))) return rc;
    return launch_row(s, st, clamp_iter(s, it));
}

// Resets the working factors from the stored W0 / H0.
int enqueue_start(nmfg_solve *s, hipStream_t st) {
    nmfg_matrix *m = s->m;
    const size_t wb = sizeof(double) * (size_t)m->n * s->R * s->batch;
    const size_t hb = sizeof(double) * (size_t)s->R * m->ldv * s->batch;
    HIPCHK(hipMemcpyAsync(s->W, s->W0, wb, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(s->H, s->H0, hb, hipMemcpyDeviceToDevice, st));
    if (s->kind == K_EHALS) HIPCHK(hipMemcpyAsync(s->Hp, s->H0, hb, hipMemcpyDeviceToDevice, st));
    if (s->kind == K_OFFSET)
        HIPCHK(hipMemcpyAsync(s->off, s->off0, sizeof(double) * (size_t)m->n * s->batch,
                              hipMemcpyDeviceToDevice, st));
    return launch_prep(s, st);
}

// Applies the pending logical transforms after `iters` iterations.
int enqueue_finish(nmfg_solve *s, hipStream_t st, int64_t iters) {
    if (iters == 0) return NMFG_OK;
    int rc = launch_wstats(s, st, 0, 0);
    if (rc) return rc;
    return launch_fin(s, st, clamp_iter(s, iters), 1);
}

int enqueue_fixed(nmfg_solve *s, hipStream_t st) {
    int rc = enqueue_start(s, st);
    if (rc) return rc;
    for (int64_t it = 1; it <= s->prm.max_iter; ++it)
        if ((rc = enqueue_iteration(s, st, it))) return rc;
    return enqueue_finish(s, st, s->prm.max_iter);
}

// Objective of the logical factors after `it` iterations, its total and the
// stop test, all on the device (k_obj partials per tile, k_objfin).
int launch_objfin_of(nmfg_solve *s, hipStream_t st, int nparts, int kl);

int objective_device(nmfg_solve *s, hipStream_t st, int64_t it) {
    // The objective reads W's statistics only where the logical W differs
    // from the stored one: lee's pending rescale, nsNMF's column
    // normalisation, dnmf's norms; every other form has divisors 1.
    const bool need_ws = (s->kind == K_LEE && s->prm.rescale) || s->kind == K_NSNMF ||
                         s->kind == K_DNMF;
    if (need_ws) {
        // phase 0 only: k_obj reduces the partials into the divisors
        int rc = launch_wstats(s, st, it == 0, 0, true);
        if (rc) return rc;
    }
    ObjArgs o{};
    nmfg_matrix *m = s->m;
    o.V = m->V;
    o.vplane = m->vplane;
    o.ldv = m->ldv;
    o.n = (int)m->n;
    o.p = (int)m->p;
    o.r = s->r;
    o.ntiles = s->ntiles_obj;
    o.osplit = s->osplit;
    o.W = s->W;
    o.H = s->H;
    o.wsc = s->wsc;
    o.unit_sc = need_ws ? 0 : 1;
    o.wpart = need_ws ? s->wpart : nullptr;
    o.wnparts = s->nparts;
    o.wkind = s->kind;
    o.wrescale = s->prm.rescale ? 1 : 0;
    o.wfirst = it == 0 ? 1 : 0;
    o.hst = s->hst;
    o.hmode = s->kind == K_DNMF     ? (it == 0 ? 0 : 1)
              : s->kind == K_BRUNET ? 2
              : s->kind == K_NSNMF  ? 3
                                    : 0;
    o.clamp_h = clamp_iter(s, it) ? 1 : 0;
    o.off = s->kind == K_OFFSET ? s->off : nullptr;
    o.theta = s->prm.theta;
    o.wgt = s->kind == K_LSNMF ? s->wgt : nullptr;
    o.wf = s->wf;
    o.eps = solve_eps(s);
    o.part = s->obj_part;
    o.kl = kl_kind(s->kind) ? 1 : 0;
    o.done = s->loop_done;
    o.mfma = s->obj_mfma ? 1 : 0;
    HIPCHK(s->K->obj(o, st));
    return launch_objfin_of(s, st, s->ntiles_obj * s->osplit, o.kl);
}

// k_objfin over `nparts` partials in obj_part (objective total, deviance
// record, stop test; iteration from the device counter).
int launch_objfin_of(nmfg_solve *s, hipStream_t st, int nparts, int kl) {
    ObjFinArgs f{};
    f.part = s->obj_part;
    f.nparts = nparts;
    f.kl = kl;
    f.it = s->d_niter + 1;
    f.stop = s->prm.stop;
    f.tol = s->prm.rel_tol;
    f.min_iter = s->prm.min_iter;
    f.interval = s->prm.check_interval > 0 ? s->prm.check_interval : 1;
    f.niter = s->prm.check_niter;
    f.th = s->prm.stationary_th;
    f.devs = s->d_devs;
    f.state = s->d_state;
    f.done = s->d_done;
    f.n_iter = s->d_niter;
    HIPCHK(launch_objfin(f, st));
    return NMFG_OK;
}

int read_done(nmfg_solve *s, hipStream_t st, int *done) {
    HIPCHK(hipMemcpyAsync(done, s->d_done, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return NMFG_OK;
}

// Loop with an objective-based stop (stationary window, or the relative
// change), decided on the device: every iteration's kernels return at once
// once the stop flag is set, and the host reads the flag every
// STOP_CHECK_CHUNK iterations, so a stop costs at most that many no-op
// iterations and no per-iteration round trip.
// Iterations between host reads of the stop flag; 10 keeps brunet's
// every-10th clamp at the same places in every chunk, so one captured graph
// serves every chunk after the first.
constexpr int STOP_CHECK_CHUNK = 10;

// Iteration `it` of a loop with an objective-based stop.  Fused (brunet): H
// saved, the column pass (which also sums f(it - 1)), the stop test of
// iteration it - 1, then the rest of the iteration; otherwise the iteration
// and its own objective.
int enqueue_step(nmfg_solve *s, hipStream_t st, int64_t it) {
    int rc;
    if (!s->fuse_obj) {
        if ((rc = enqueue_iteration(s, st, it))) return rc;
        return objective_device(s, st, it);
    }
    HIPCHK(launch_copy_unless_done(s->H, s->Hs, (int64_t)s->R * s->m->ldv, s->loop_done, st));
    if ((rc = launch_col(s, st, it == 1, clamp_iter(s, it - 1)))) return rc;
    if ((rc = launch_objfin_of(s, st, s->ntiles_col * s->kl_split, 1))) return rc;
    if ((rc = launch_red(s, st, s->ntiles_col))) return rc;
    return launch_row(s, st, clamp_iter(s, it));
}

// One chunk of STOP_CHECK_CHUNK iterations with their objectives, iterations
// iter + 1 .. iter + STOP_CHECK_CHUNK (iter a multiple of the chunk, past the
// first chunk), as a graph captured on first use and kept in `slot`.
// NMFG_EUNSUPPORTED when capture is not possible (the caller then enqueues
// the chunk directly).
int launch_chunk_graph(nmfg_solve *s, hipStream_t st, int64_t iter, ChunkGraph *slot) {
    if (slot->failed || s->no_graph) return NMFG_EUNSUPPORTED;
    if (!slot->exec) {
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            slot->failed = true;
            return NMFG_EUNSUPPORTED;
        }
        int rc = NMFG_OK;
        for (int64_t it = iter + 1; it <= iter + STOP_CHECK_CHUNK && !rc; ++it)
            rc = enqueue_step(s, st, it);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(st, &graph);
        if (rc || e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            slot->failed = true;
            (void)hipGetLastError();
            return rc ? rc : NMFG_EUNSUPPORTED;
        }
        e = hipGraphInstantiate(&slot->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            slot->exec = nullptr;
            slot->failed = true;
            (void)hipGetLastError();
            return NMFG_EUNSUPPORTED;
        }
    }
    HIPCHK(hipGraphLaunch(slot->exec, st));
    return NMFG_OK;
}

// The solve's own chunk graph (single-solve runs, on the context stream).
int launch_chunk_graph(nmfg_solve *s, hipStream_t st, int64_t iter) {
    ChunkGraph slot{s->chunk_graph, s->chunk_graph_failed};
    const int rc = launch_chunk_graph(s, st, iter, &slot);
    s->chunk_graph = slot.exec;
    s->chunk_graph_failed = slot.failed;
    return rc;
}

// The pieces of a loop with an objective-based stop, shared by a single
// solve (run_with_objective) and a step group (nmfg_group_step, several
// solves advanced together).  While a run is armed, s->loop_done points at
// the solve's stop flag (every kernel of the loop returns at once once it is
// set) and s->fused_active says whether brunet's objective rides in its
// column pass (f(0) from the first column pass, f(t - 1) from the t-th).
bool objrun_fused(const nmfg_solve *s) { return s->fuse_obj && s->prm.max_iter > 0; }

void objrun_disarm(nmfg_solve *s) {
    s->loop_done = nullptr;
    s->fused_active = false;
}

// Start: stop state reset (window 0, flag 0, stop iteration -1, evaluation
// counter 0), the working factors from W0 / H0 and, unfused, f(0).
int objrun_begin(nmfg_solve *s, hipStream_t st) {
    s->devs.clear();
    HIPCHK(launch_stop_reset(s->d_state, s->d_done, s->d_niter, st));
    int rc = enqueue_start(s, st);
    if (rc) return rc;
    s->loop_done = s->d_done;
    s->fused_active = false;
    if (!objrun_fused(s) && (rc = objective_device(s, st, 0))) return rc;
    s->fused_active = objrun_fused(s);
    return NMFG_OK;
}

// Iterations *iter + 1 .. the end of the next chunk; a full chunk past the
// first replays a captured chunk graph (the kernels of every chunk after
// the first are the same): the solve's own, or, for a step group, the one
// in *slot, which the group captures on its own stream and keeps (a graph
// never outlives the streams it is launched on).
int objrun_chunk(nmfg_solve *s, hipStream_t st, int64_t *iter, ChunkGraph *slot = nullptr) {
    const int64_t chunk_end = std::min<int64_t>(s->prm.max_iter, *iter + STOP_CHECK_CHUNK);
    int rc;
    if (*iter >= STOP_CHECK_CHUNK && chunk_end - *iter == STOP_CHECK_CHUNK &&
        (rc = slot ? launch_chunk_graph(s, st, *iter, slot)
                   : launch_chunk_graph(s, st, *iter)) != NMFG_EUNSUPPORTED) {
        if (rc) return rc;
        *iter = chunk_end;
        return NMFG_OK;
    }
    while (*iter < chunk_end) {
        *iter += 1;
        if ((rc = enqueue_step(s, st, *iter))) return rc;
    }
    return NMFG_OK;
}

// End of a run after `iter` iterations, `done` the stop flag and `stop_it`
// the device's stop iteration: the factors of the stopping iteration, the
// deviances (s->devs, copied on st: valid once st is synchronised), n_iter,
// the pending logical transforms.  Disarms the run.
int objrun_finish(nmfg_solve *s, hipStream_t st, int64_t iter, int done, int64_t stop_it) {
    const bool fused = s->fused_active;
    objrun_disarm(s);
    int rc;
    if (done && !fused && stop_it == 0) {
        // stopped at f(0): no iteration counts (the loop's kernels did nothing)
        s->n_iter = 0;
        s->devs.resize(1);
        HIPCHK(hipMemcpyAsync(s->devs.data(), s->d_devs, sizeof(double), hipMemcpyDeviceToHost, st));
        if (s->kind == K_DNMF) {
            // early return normalises W and H without folding d (lib.rs:849-853)
            if ((rc = launch_hstats(s, st)) || (rc = launch_wstats(s, st, 1, 1))) return rc;
            return launch_fin(s, st, 0, 0);
        }
        return NMFG_OK;
    }
    if (fused) {
        if (done) {
            // stopped at iteration i < iter: its H, kept before pass i + 1
            HIPCHK(hipMemcpyAsync(s->H, s->Hs, sizeof(double) * (size_t)s->R * s->m->ldv,
                                  hipMemcpyDeviceToDevice, st));
        } else if ((rc = objective_device(s, st, iter))) {
            return rc;  // f(max_iter): no later pass computes it
        }
    }
    const int64_t n_iter = done ? stop_it : iter;
    s->n_iter = n_iter;
    s->devs.resize((size_t)n_iter + 1);
    HIPCHK(hipMemcpyAsync(s->devs.data(), s->d_devs, sizeof(double) * (n_iter + 1),
                          hipMemcpyDeviceToHost, st));
    return enqueue_finish(s, st, n_iter);
}

int run_with_objective(nmfg_solve *s, hipStream_t st) {
    struct Disarm {
        nmfg_solve *s;
        ~Disarm() { objrun_disarm(s); }
    } disarm{s};
    int rc = objrun_begin(s, st);
    if (rc) return rc;
    int done = 0;
    int64_t iter = 0;
    if (!objrun_fused(s) && (rc = read_done(s, st, &done))) return rc;
    while (!done && iter < s->prm.max_iter) {
        if ((rc = objrun_chunk(s, st, &iter))) return rc;
        if ((rc = read_done(s, st, &done))) return rc;
    }
    int64_t stop_it = -1;
    if (done) {
        HIPCHK(hipMemcpyAsync(&stop_it, s->d_niter, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    rc = objrun_finish(s, st, iter, done, stop_it);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(st));
    return NMFG_OK;
}

int validate_params(const nmfg_params *prm) {
    if (!prm) return fail(NMFG_EINVAL, "params is NULL");
    if (kind_of(prm->algo) < 0) {
        if (prm->algo >= NMFG_BRUNET && prm->algo <= NMFG_SNMF_L)
            return fail(NMFG_EUNSUPPORTED,
                        "method not implemented on the GPU path: supported brunet, lee, offset, "
                        "nsNMF, hals, ehals, dnmf");
        return fail(NMFG_EINVAL, "unknown algorithm code");
    }
    if (prm->stop != NMFG_STOP_MAX_ITER && prm->stop != NMFG_STOP_STATIONARY &&
        prm->stop != NMFG_STOP_REL_CHANGE)
        return fail(NMFG_EINVAL, "unknown stopping criterion");
    if (prm->stop == NMFG_STOP_REL_CHANGE && !(prm->rel_tol >= 0.0))
        return fail(NMFG_EINVAL, "rel_tol must be >= 0");
    if (prm->max_iter < 0) return fail(NMFG_EINVAL, "max_iter must be >= 0");
    if (!valid_form(prm->precision)) return fail(NMFG_EINVAL, "unknown precision code");
    return NMFG_OK;
}

// Host factors (batch consecutive pairs of n x r and r x p, row-major) to
// the device layout (n x R and R x ldv per solve, zero padded) and back.
// Rank of batch member b, and the offsets (in doubles) of member b's W and
// H in host buffers holding the members' n x r_b and r_b x p factors one
// after another.
int member_rank(const nmfg_solve *s, int b) { return s->mranks.empty() ? s->r : s->mranks[b]; }
size_t member_offset(const nmfg_solve *s, int b) {
    size_t o = 0;
    for (int c = 0; c < b; ++c) o += (size_t)member_rank(s, c);
    return o;
}

int pack_factors(nmfg_solve *s, const double *W0, const double *H0, double *dW, double *dH) {
    nmfg_matrix *m = s->m;
    const int R = s->R;
    const size_t ws = (size_t)m->n * R, hs = (size_t)R * m->ldv;
    std::vector<double> w(ws * s->batch, 0.0), h(hs * s->batch, 0.0);
    for (int b = 0; b < s->batch; ++b) {
        const int r = member_rank(s, b);
        const size_t ob = member_offset(s, b);
        const double *W0b = W0 ? W0 + ob * m->n : nullptr;
        const double *H0b = H0 ? H0 + ob * m->p : nullptr;
        if (W0b)
            for (int64_t i = 0; i < m->n; ++i)
                for (int k = 0; k < r; ++k) w[b * ws + (size_t)i * R + k] = W0b[i * r + k];
        if (H0b)
            for (int k = 0; k < r; ++k)
                std::memcpy(&h[b * hs + (size_t)k * m->ldv], H0b + (int64_t)k * m->p,
                            sizeof(double) * m->p);
    }
    if (dW)
        HIPCHK(hipMemcpyAsync(dW, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice,
                              m->ctx->stream));
    if (dH)
        HIPCHK(hipMemcpyAsync(dH, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                              m->ctx->stream));
    HIPCHK(hipStreamSynchronize(m->ctx->stream));
    return NMFG_OK;
}

int unpack_factors(nmfg_solve *s, const double *dW, const double *dH, double *W, double *H) {
    nmfg_matrix *m = s->m;
    const int R = s->R;
    const size_t ws = (size_t)m->n * R, hs = (size_t)R * m->ldv;
    std::vector<double> w(ws * s->batch), h(hs * s->batch);
    HIPCHK(hipMemcpyAsync(w.data(), dW, w.size() * sizeof(double), hipMemcpyDeviceToHost,
                          m->ctx->stream));
    HIPCHK(hipMemcpyAsync(h.data(), dH, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                          m->ctx->stream));
    HIPCHK(hipStreamSynchronize(m->ctx->stream));
    for (int b = 0; b < s->batch; ++b) {
        const int r = member_rank(s, b);
        const size_t ob = member_offset(s, b);
        if (W)
            for (int64_t i = 0; i < m->n; ++i)
                for (int k = 0; k < r; ++k)
                    W[ob * m->n + i * r + k] = w[b * ws + (size_t)i * R + k];
        if (H)
            for (int k = 0; k < r; ++k)
                std::memcpy(H + ob * m->p + (int64_t)k * m->p, &h[b * hs + (size_t)k * m->ldv],
                            sizeof(double) * m->p);
    }
    return NMFG_OK;
}

// Rows per work-group of the lsNMF row pass (two accumulators per entry).
int ls_rows_of(int R) { return ls_rows(R); }

// Sizes, tiling and split choices of a solve (no device memory yet).
int solve_setup(nmfg_matrix *m, int32_t r, const nmfg_params *prm, int kind, int batch,
                nmfg_solve *s) {
    if (r < 1) return fail(NMFG_EINVAL, "rank must be >= 1");
    if (batch < 1 || batch > 65535) return fail(NMFG_EINVAL, "batch must be in 1..65535");
    if (batch > 1 && prm->stop != NMFG_STOP_MAX_ITER)
        return fail(NMFG_EUNSUPPORTED, "a batched solve runs a fixed number of iterations");
    const RankKernels *K = kernels_for_rank(r, m->precision);
    if (!K)
        return fail(NMFG_EUNSUPPORTED, "rank " + std::to_string(r) +
                                           " exceeds the supported maximum " +
                                           std::to_string(max_supported_rank()));
    s->m = m;
    s->prm = *prm;
    s->r = r;
    s->R = K->R;
    s->K = K;
    s->kind = kind;
    s->batch = batch;
    const int R = s->R;
    const int64_t n = m->n, p = m->p;
    // FP32 form: the multiplicative updates (lee, brunet) take their
    // products in FP32.  Their updates multiply by ratios of the products,
    // so a product's relative error stays a relative error of the update.
    // The HALS family subtracts G H from W^T V (and K W from V H^T), where
    // the two nearly cancel: FP32 accumulation errors (~sqrt(n) 2^-24 of
    // the product) become large errors of the difference, flip eps clamps
    // and send the iterates elsewhere (2000 x 8000 single-cell, rank 12:
    // max-relative W difference 4 after 5 iterations, where FP64 on the
    // float-rounded V stays within 1e-6).  Those kinds, and the others, run
    // FP64 arithmetic on the float V.
    s->f32 = m->form == NMFG_FP32 && (kind == K_LEE || kind == K_BRUNET);
    // BF16 form: lee's two products; the other methods run FP64 arithmetic
    // on the hi + lo planes (the HALS family for the reason above; brunet's
    // products are formed per element, W H inside V / (W H), on the VALU)
    s->b16 = m->form == NMFG_BF16X2 && kind == K_LEE;
    // lee's low-precision MFMA passes (FP32 or BF16 form): Gram from
    // k_wgram, numerator pass + k_colfin_f32, MFMA row pass + k_rowfin
    const bool lp = (s->f32 && !kl_kind(kind)) || s->b16;
    const int tw = col_tile_width(kind, R, s->f32 || s->b16);
    s->ntiles_col = (int)((p + tw - 1) / tw);
    s->ntiles_h = (int)((p + TJC - 1) / TJC);
    s->ntiles_obj = (int)((p + TJO - 1) / TJO);
    {
        // Grid coverage, checked on the host before any launch: each
        // column-tiled pass's tiles cover the p columns and its last tile
        // ends inside V's padded rows (ldv).  A grid sized for another tile
        // width than its kernel indexes with would read past V, W or H.
        const int64_t ldv = m->ldv;
        auto covers = [&](int64_t tiles, int64_t width) {
            return tiles * width >= p && tiles * width <= ldv;
        };
        bool ok = covers(s->ntiles_col, tw) && covers(s->ntiles_h, TJC) && covers(s->ntiles_obj, TJO);
        // lee's FP32 numerator pass: (ntiles_col + 1) / 2 work-groups of TJS columns
        if (s->f32 && !kl_kind(kind)) ok = ok && covers((s->ntiles_col + 1) / 2, TJS);
        if (!ok) return fail(NMFG_EINVAL, "internal: a column pass's grid does not cover V's columns");
    }
    // objective row splits: about 12 work-groups of NWO waves per CU, and at
    // least 16 rows per wave for the bandwidth-bound Frobenius terms, 8 for
    // the KL terms (bound by the division and log: more work-groups help at
    // small sizes, 500 x 1000 brunet 48.5 -> 41.5 us per iteration with the
    // relative-change stop)
    s->osplit = (int)std::max<int64_t>(
        1, std::min<int64_t>(3072 / std::max(1, s->ntiles_obj), n / (NWO * (kl_kind(kind) ? 8 : 16))));
    // W H of the Frobenius-type objectives on the FP64 matrix cores from
    // rank 8 on (2000 x 7750 rank 10: 43 -> 39 us per evaluation; 1500 x
    // 80000 rank 24 on float V: 268 -> 187; 2400 x 8400 rank 30: 110 -> 54);
    // the KL terms stay on the VALU loop, bound by the division and log
    // (1200 x 80000 rank 22: 460 -> 517 us on the matrix cores)
    s->obj_mfma = R >= 8 && !kl_kind(kind);
    if (const char *e = std::getenv("NMFG_OBJ_MFMA")) s->obj_mfma = std::atoi(e) != 0;
    if (s->obj_mfma) {
        // MFMA objective: about two work-groups per CU, at least 4 row blocks
        // of 16 per wave (each work-group first stages its H columns)
        int64_t wg = 512;
        if (const char *e = std::getenv("NMFG_OBJ_WG")) wg = std::max(1, std::atoi(e));
        s->osplit = (int)std::max<int64_t>(
            1, std::min<int64_t>(wg / std::max(1, s->ntiles_obj), n / (NWO * 64)));
    }
    const int rw = kind == K_LSNMF ? ls_rows_of(R)
                   : kl_kind(kind)  ? kl_rows_per_rowblk(R)
                                    : rows_per_rowblk(R);
    s->nrowblk = (int)((n + rw - 1) / rw);
    s->nparts = (int)((n + WSTAT_ROWS - 1) / WSTAT_ROWS);
    s->ngparts = (int)((n + WGRAM_ROWS - 1) / WGRAM_ROWS);
    // offset always takes W's record from k_wgram: its third section carries
    // W^T off there
    // External Gram: at large p (many tiles would each form it), and above
    // rank 16, where forming it in every tile doubles the column pass's MFMA
    // work (measured at 2000 x 7750: rank 20 lee 9688 -> 10133 it/s; below
    // rank 16 the two extra launches cost more than they save).
    s->xgram = kind == K_OFFSET || lp ||
               (!kl_kind(kind) && kind != K_PROD && kind != K_LSNMF &&
                (s->ntiles_col > XGRAM_TILES || R > 16));
    if (const char *e = std::getenv("NMFG_XGRAM"))
        s->xgram = kind == K_OFFSET || lp ||
                   (std::atoi(e) != 0 && !kl_kind(kind) && kind != K_PROD && kind != K_LSNMF);
    // (lsNMF has its own row pass with its own split rule)
    s->rowmfma = (R >= ROW_MFMA_MIN_R || lp) && !kl_kind(kind) && kind != K_LSNMF;
    if (const char *e = std::getenv("NMFG_ROW_MFMA"))
        s->rowmfma = lp || (std::atoi(e) != 0 && !kl_kind(kind) && kind != K_LSNMF);
    // VALU row pass: one wave per row block once the blocks (times the
    // batch) are at least one per SIMD -- longer loops, more waves resident
    s->row_waves = (int64_t)s->nrowblk * batch >= ROW_ONE_WAVE_BLOCKS ? 1 : NWR;
    if (const char *e = std::getenv("NMFG_ROW_WAVES")) s->row_waves = std::atoi(e) == 1 ? 1 : NWR;
    if (kind == K_LSNMF) {
        // lsNMF passes: the same split rules as the KL passes below
        const int64_t cus = 256;
        const int64_t wgs = (int64_t)s->ntiles_col * batch;
        const int64_t slots = cus * K->ls_occupancy(0);
        const int64_t nwk = R <= 6 ? 16 : 8;
        if (wgs < slots)
            s->kl_split = (int)std::max<int64_t>(1, std::min<int64_t>(slots / wgs, n / (nwk * 32)));
        const int64_t rwgs = (int64_t)s->nrowblk * batch;
        const int64_t rslots = cus * K->ls_occupancy(1);
        if (rwgs < rslots)
            s->nsplit = (int)std::max<int64_t>(1, std::min<int64_t>(rslots / rwgs, m->ldv / 64 / (NWR * 4)));
        if (const char *e = std::getenv("NMFG_KL_SPLIT")) s->kl_split = std::max(1, std::atoi(e));
        if (const char *e = std::getenv("NMFG_KL_ROW_SPLIT")) s->nsplit = std::max(1, std::atoi(e));
    }
    if (kl_kind(kind)) {
        // KL passes: split the rows of the column pass / the columns of the
        // row pass when their work-groups (times the batch) would leave CU
        // slots idle, up to one full round of resident work-groups
        // (occupancy from the compiled kernels); at least 32 rows per wave
        // and 4 column steps per wave.
        const int64_t cus = 256;
        const int64_t wgs = (int64_t)s->ntiles_col * batch;
        const int64_t slots = cus * K->kl_occupancy(0, s->f32 ? 1 : 0);
        const int64_t nwk = R <= 6 ? 16 : 8;
        if (wgs < slots)
            s->kl_split = (int)std::max<int64_t>(1, std::min<int64_t>(slots / wgs, n / (nwk * 32)));
        if (const char *e = std::getenv("NMFG_KL_SPLIT")) s->kl_split = std::max(1, std::atoi(e));
        const int64_t rwgs = (int64_t)s->nrowblk * batch;
        const int64_t rslots = cus * K->kl_occupancy(1, s->f32 ? 1 : 0);
        const int64_t rsteps = m->ldv / (64 * (R <= 16 ? 2 : 1));  // column steps of k_row
        if (s->row_waves != 1 && rwgs < rslots)
            s->nsplit = (int)std::max<int64_t>(1, std::min<int64_t>(rslots / rwgs, rsteps / (NWR * 4)));
        if (const char *e = std::getenv("NMFG_KL_ROW_SPLIT")) s->nsplit = std::max(1, std::atoi(e));
    }
    // brunet with an objective-based stop sums the objective inside the next
    // KL column pass when that pass has a work-group per CU or more (on fewer
    // work-groups the added log terms run serially on a few CUs: 500 x 1000
    // rank 5 42.7 -> 56.1 us per iteration, while 2000 x 7750 rank 10 gains
    // 155.7 -> 132.1)
    s->fuse_obj = kind == K_BRUNET && prm->stop != NMFG_STOP_MAX_ITER && batch == 1 && !s->f32 &&
                  (int64_t)s->ntiles_col * s->kl_split >= 256;
    if (const char *e = std::getenv("NMFG_FUSE_OBJ")) s->fuse_obj = s->fuse_obj && std::atoi(e) != 0;
    if (s->rowmfma) {
        // 16-row tiles per block: each H' load feeds this many row tiles.
        // Measured: at p = 500000, rank 20, 4 tiles 2163 us vs 3396 us for 1;
        // at p = 7750, rank 20, 2 tiles 55.8 us vs 58.1 (4 tiles 68.3: too
        // few columns per work-group); no difference at ranks <= 16.
        s->row_tiles = p >= 32768 ? 4 : (R > 16 ? 2 : 1);
        // FP32 form: a work-group is 4 waves of row_tiles 16-row tiles
        // sharing the staged H' chunk
        if (lp) s->row_tiles = n >= 1024 ? 4 : (n >= 256 ? 2 : 1);
        if (const char *e = std::getenv("NMFG_ROW_TILES")) {
            const int t = std::atoi(e);
            s->row_tiles = t == 4 ? 4 : (t == 2 ? 2 : 1);
        }
        const int64_t rows_blk = (int64_t)RB_MFMA * s->row_tiles * (lp ? 4 : 1);
        s->nrb = (int)((n + rows_blk - 1) / rows_blk);
        // column splits: the fewest that minimise rounds of resident
        // work-groups per split (time ~ ceil(nrb x nsplit x batch / slots) /
        // nsplit); a count just above a full round would leave a second,
        // nearly empty one (5000 x 100000 rank 20: 553 work-groups on 512
        // slots); searched up to twice the splits of one full round (more
        // only trade partial-sum traffic for finer rounding)
        // column steps of a split (FP32 form: 64-column chunks, at least 8 each)
        const int64_t steps = lp ? m->ldv / 64 * 8 : m->ldv / 8;
        const int64_t slots = 256 * (int64_t)K->rowmfma_occupancy(s->row_tiles, lp ? 1 : 0);
        const int64_t blocks = (int64_t)s->nrb * batch;
        const int64_t max_split =
            std::max<int64_t>(1, std::min<int64_t>(steps / 8, 2 * ((slots + blocks - 1) / blocks)));
        // The finishing kernel reads every split's partial per row, which
        // the round model leaves out: take the fewest splits within 10 % of
        // the best cost (1500 x 10000 rank 16: 10 splits rather than 21,
        // k_rowfin 27 -> a few us).
        double best_cost = 1e300;
        for (int64_t ns = 1; ns <= max_split; ++ns)
            best_cost = std::min(best_cost, (double)((blocks * ns + slots - 1) / slots) / (double)ns);
        int64_t best = max_split;
        for (int64_t ns = 1; ns <= max_split; ++ns)
            if ((double)((blocks * ns + slots - 1) / slots) / (double)ns <= 1.1 * best_cost) {
                best = ns;
                break;
            }
        s->nsplit = (int)best;
        if (const char *e = std::getenv("NMFG_ROW_SPLIT")) s->nsplit = std::max(1, std::atoi(e));
    }
    if (s->b16) {
        // BF16 numerator pass: split the 32-row steps across work-groups
        // when the tiles (times the batch) leave resident work-group slots
        // idle, up to one full round, with at least 8 steps per wave
        const int64_t wgs = (int64_t)s->ntiles_col * batch;
        const int64_t slots = 256 * (int64_t)K->kl_occupancy(2, 1);
        const int64_t steps = b16_rows(n) / 32;
        if (wgs < slots)
            s->kl_split = (int)std::max<int64_t>(1, std::min<int64_t>(slots / wgs, steps / (4 * 8)));
        if (const char *e = std::getenv("NMFG_KL_SPLIT")) s->kl_split = std::max(1, std::atoi(e));
    }
    return NMFG_OK;
}

// The solve's device buffers as 256-byte aligned pieces of one block at
// `base`; returns the block's size (base == nullptr: the size only).
size_t solve_bind(nmfg_solve *s, char *base) {
    nmfg_matrix *m = s->m;
    const int R = s->R, kind = s->kind, batch = s->batch;
    const int64_t n = m->n;
    const nmfg_params *prm = &s->prm;
    const size_t B = (size_t)batch;
    const size_t wsz = sizeof(double) * (size_t)n * R * B;
    const size_t hsz = sizeof(double) * (size_t)R * m->ldv * B;
    const int nrec = nrec_of(s);
    struct {
        void **p;
        size_t b;
    } req[] = {
        {(void **)&s->W0, wsz},
        {(void **)&s->H0, hsz},
        {(void **)&s->W, wsz},
        {(void **)&s->Wx, kind == K_EHALS || kind == K_NSNMF ? wsz : 8},
        {(void **)&s->H, hsz},
        {(void **)&s->Hp, kind == K_EHALS || kind == K_NSNMF ? hsz : 8},
        {(void **)&s->off0, kind == K_OFFSET ? sizeof(double) * (size_t)n * B : 8},
        {(void **)&s->off, kind == K_OFFSET ? sizeof(double) * (size_t)n * B : 8},
        {(void **)&s->rsv, kind == K_OFFSET ? sizeof(double) * (size_t)n : 8},
        {(void **)&s->red, sizeof(double) * (size_t)nrec * nred(R) * B},
        {(void **)&s->hst, sizeof(double) * (size_t)nred(R) * B},
        {(void **)&s->wsx, sizeof(double) * (size_t)R * B},
        {(void **)&s->wpart, sizeof(double) * (size_t)s->nparts * 2 * R * B},
        {(void **)&s->wsc, sizeof(double) * (size_t)3 * R * B},
        {(void **)&s->obj_part, sizeof(double) * (size_t)std::max<int64_t>((int64_t)s->ntiles_obj * s->osplit,
                                                                          (int64_t)s->ntiles_col * s->kl_split)},
        {(void **)&s->Hs, s->fuse_obj ? hsz : 8},
        {(void **)&s->gpart, s->xgram ? sizeof(double) * (size_t)s->ngparts * nred(R) * B : 8},
        {(void **)&s->gram, s->xgram ? sizeof(double) * (size_t)nred(R) * B : 8},
        {(void **)&s->rpart, s->rowmfma || s->nsplit > 1 ? sizeof(double) * (size_t)batch_of(s).rpart * B : 8},
        {(void **)&s->kpart, s->kl_split > 1 && !s->b16 ? sizeof(double) * (size_t)batch_of(s).kpart * B : 8},
        {(void **)&s->d_done, sizeof(int)},
        {(void **)&s->d_niter, sizeof(int64_t) * 2},
        {(void **)&s->d_devs, sizeof(double) * (size_t)(prm->stop != NMFG_STOP_MAX_ITER ? prm->max_iter + 1 : 1)},
        {(void **)&s->d_state, sizeof(double) * 4},
        {(void **)&s->Wf, s->f32 ? sizeof(float) * (size_t)n * R * B : 8},
        {(void **)&s->Hf, s->f32 ? sizeof(float) * (size_t)R * m->ldv * B : 8},
        {(void **)&s->num, (s->f32 && !kl_kind(kind)) || s->b16 ? sizeof(float) * (size_t)batch_of(s).num * B : 8},
        {(void **)&s->Wb, s->b16 ? sizeof(uint16_t) * (size_t)batch_of(s).wb * B : 8},
        {(void **)&s->Hb, s->b16 ? sizeof(uint16_t) * (size_t)batch_of(s).hb * B : 8},
        {(void **)&s->d_ranks, sizeof(int) * B},
    };
    size_t total = 0;
    for (auto &q : req) {
        if (base) *q.p = base + total;
        total += align256(std::max<size_t>(q.b, 8));
    }
    return total;
}

// Deterministic contents for buffers a kernel may read before writing,
// filled on the device (no host copies).
hipError_t solve_init_buffers(nmfg_solve *s) {
    nmfg_matrix *m = s->m;
    const int R = s->R, kind = s->kind;
    const int64_t B = s->batch;
    hipStream_t st = m->ctx->stream;
    hipError_t e = launch_fill_f64(s->hst, (int64_t)nred(R) * B, 0.0, st);
    if (e == hipSuccess) e = launch_fill_f64(s->wsc, (int64_t)3 * R * B, 0.0, st);
    if (e == hipSuccess && (kind == K_EHALS || kind == K_NSNMF))
        e = launch_fill_f64(s->Wx, m->n * R * B, 0.0, st);
    // nsNMF: the column pass writes S H' for columns j < p only; the row
    // pass reads Hp's padding columns p .. ldv - 1 (against V's zero
    // padding), which must hold zeros, not what the block held before.
    if (e == hipSuccess && kind == K_NSNMF) e = launch_fill_f64(s->Hp, (int64_t)R * m->ldv * B, 0.0, st);
    // BF16 form: k_wgram writes Wb's rows i < n of columns k < R only, the
    // column pass writes Hb's columns j < p only; the rest is read as zeros
    if (e == hipSuccess && s->b16)
        e = launch_fill_f64(reinterpret_cast<double *>(s->Wb), (int64_t)batch_of(s).wb * B / 4, 0.0, st);
    if (e == hipSuccess && s->b16)
        e = launch_fill_f64(reinterpret_cast<double *>(s->Hb), (int64_t)batch_of(s).hb * B / 4, 0.0, st);
    // FP32 form: the column pass writes H's float copy for j < p only; the
    // row pass reads its padding columns (against V's zero padding)
    if (e == hipSuccess && s->f32)
        e = launch_fill_f64(reinterpret_cast<double *>(s->Hf), (int64_t)R * m->ldv * B / 2, 0.0, st);
    return e;
}

// Choices a caller (the sweep's autotuner) may impose on a solve instead of
// the occupancy rules: 0 keeps the rule's value.
struct SolveOverrides {
    int kl_split = 0;  // KL column pass row splits
    int nsplit = 0;    // row pass column splits
};

int solve_create_kind(nmfg_matrix *m, int32_t r, const nmfg_params *prm, int kind, int batch,
                      nmfg_solve **out, const int *member_ranks = nullptr,
                      const SolveOverrides *ov = nullptr) {
    nmfg_solve *s = new nmfg_solve();
    int rc = solve_setup(m, r, prm, kind, batch, s);
    if (rc) {
        delete s;
        return rc;
    }
    if (member_ranks) s->mranks.assign(member_ranks, member_ranks + batch);
    // (row splits of the column pass exist in the KL passes; column splits
    // of the row pass in the KL and MFMA row passes)
    if (ov && ov->kl_split > 0 && kl_kind(kind)) s->kl_split = ov->kl_split;
    if (ov && ov->nsplit > 0 && (kl_kind(kind) || s->rowmfma)) s->nsplit = ov->nsplit;
    // one block for all buffers, each piece 256-byte aligned
    void *blk = nullptr;
    if ((rc = dmalloc(s, &blk, solve_bind(s, nullptr)))) {
        solve_free(s);
        delete s;
        return rc;
    }
    solve_bind(s, static_cast<char *>(blk));
    const int64_t n = m->n, p = m->p;
    const size_t B = (size_t)batch;
    hipError_t e = solve_init_buffers(s);
    if (e == hipSuccess && !s->mranks.empty())
        e = hipMemcpyAsync(s->d_ranks, s->mranks.data(), sizeof(int) * B, hipMemcpyHostToDevice,
                           m->ctx->stream);
    if (e == hipSuccess && kind == K_OFFSET) {
        // row sums of V, and the default initial offset rowMeans(V)
        // (lib.rs:1669-1683) for every solve of the batch
        e = launch_rowsum_v(m->precision, m->V, m->vplane, m->ldv, (int)n, (int)p, s->rsv,
                            m->ctx->stream);
        std::vector<double> rs((size_t)n);
        if (e == hipSuccess)
            e = hipMemcpyAsync(rs.data(), s->rsv, sizeof(double) * n, hipMemcpyDeviceToHost,
                               m->ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(m->ctx->stream);
        std::vector<double> o0((size_t)n * B);
        for (size_t b = 0; b < B; ++b)
            for (int64_t i = 0; i < n; ++i) o0[b * n + i] = rs[i] / (double)p;
        if (e == hipSuccess)
            e = hipMemcpyAsync(s->off0, o0.data(), sizeof(double) * o0.size(), hipMemcpyHostToDevice,
                               m->ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(m->ctx->stream);
    if (e != hipSuccess) {
        solve_free(s);
        delete s;
        return fail(NMFG_EDEVICE, std::string("init: ") + hipGetErrorString(e));
    }
    *out = s;
    return NMFG_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

void nmfg_params_init(nmfg_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->algo = NMFG_BRUNET;
    p->max_iter = 2000;
    p->eps = 1e-9;
    p->rescale = 1;
    p->theta = 0.5;
    p->stop = NMFG_STOP_MAX_ITER;
    p->stationary_th = 2.220446049250313e-16;
    p->check_interval = 50;
    p->check_niter = 10;
    p->num_threads = 0;
    p->precision = NMFG_FP64;
    p->rel_tol = 1e-6;
    p->min_iter = 1;
}

const char *nmfg_last_error(void) { return g_err.c_str(); }

int32_t nmfg_max_rank(void) { return max_supported_rank(); }

int nmfg_set_num_threads(int64_t n) {
    static std::once_flag once;
    static bool first = false;
    (void)n;
    std::call_once(once, [] { first = true; });
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    const bool r = first;
    first = false;
    return r ? 1 : 0;
}

int nmfg_smoothing_matrix(int32_t r, double theta, double *S) {
    if (r < 0 || (!S && r > 0)) return fail(NMFG_EINVAL, "bad smoothing matrix arguments");
    const double off = theta / (double)r;
    const double diag = (1.0 - theta) + off;
    for (int i = 0; i < r; ++i)
        for (int j = 0; j < r; ++j) S[i * r + j] = i == j ? diag : off;
    return NMFG_OK;
}

int nmfg_ctx_create(int32_t device, nmfg_ctx **out) {
    if (!out) return fail(NMFG_EINVAL, "out is NULL");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0)
        return fail(NMFG_EDEVICE, "no HIP device available (the library has no CPU fallback)");
    if (device < 0 || device >= ndev) return fail(NMFG_EINVAL, "device index out of range");
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(NMFG_EDEVICE, std::string("device is ") + prop.gcnArchName +
                                      "; this library is built for gfx950 only");
    if (device < kMaxDevices) {
        // once per device, with the device current: this device's copy of
        // the no-allocation scratch array
        NoallocDevice &nd = g_noalloc[device];
        std::call_once(nd.once, [&nd] { nd.err = noalloc_scratch_address(&nd.ptr); });
    }
    nmfg_ctx *c = new nmfg_ctx();
    c->device = device;
    c->idle_cap = (size_t)std::max(0, env_int("NMFG_CACHE_MB", 8192)) << 20;
    c->poison = env_int("NMFG_POISON", 0) != 0;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(NMFG_EDEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    *out = c;
    return NMFG_OK;
}

int nmfg_ctx_destroy(nmfg_ctx *c) {
    if (!c) return NMFG_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    {
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        if (c->budget) {
            // blocks still held (matrices or solves not destroyed) stay charged
            std::lock_guard<std::mutex> ll(c->budget->list_mu);
            auto &v = c->budget->ctxs;
            v.erase(std::remove(v.begin(), v.end(), c), v.end());
        }
        ctx_trim(c);
    }
    (void)hipStreamDestroy(c->stream);
    delete c;
    return NMFG_OK;
}

int nmfg_ctx_release_cached(nmfg_ctx *c) {
    if (!c) return fail(NMFG_EINVAL, "NULL argument");
    Guard g(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    ctx_trim(c);
    return NMFG_OK;
}

void *nmfg_ctx_stream(nmfg_ctx *c) { return c ? (void *)c->stream : nullptr; }

int nmfg_budget_create(int64_t bytes, nmfg_budget **out) {
    if (!out) return fail(NMFG_EINVAL, "out is NULL");
    if (bytes < 0) return fail(NMFG_EINVAL, "budget must be >= 0 bytes");
    nmfg_budget *b = new nmfg_budget();
    b->cap = bytes;
    *out = b;
    return NMFG_OK;
}

int nmfg_budget_destroy(nmfg_budget *b) {
    if (!b) return NMFG_OK;
    {
        std::lock_guard<std::mutex> lk(b->list_mu);
        if (!b->ctxs.empty()) return fail(NMFG_EINVAL, "contexts are still attached to the budget");
    }
    delete b;
    return NMFG_OK;
}

int nmfg_ctx_set_budget(nmfg_ctx *c, nmfg_budget *b) {
    if (!c) return fail(NMFG_EINVAL, "NULL context");
    Guard g(c);
    auto held = [c] {
        int64_t h = 0;
        for (auto &kv : c->block_bytes) h += (int64_t)kv.second;
        return h;
    };
    if (nmfg_budget *old = c->budget) {
        std::lock_guard<std::mutex> ll(old->list_mu);
        std::lock_guard<std::mutex> pl(c->pool);
        budget_charge(old, -held());
        old->ctxs.erase(std::remove(old->ctxs.begin(), old->ctxs.end(), c), old->ctxs.end());
        c->budget = nullptr;
    }
    if (b) {
        std::lock_guard<std::mutex> ll(b->list_mu);
        std::lock_guard<std::mutex> pl(c->pool);
        budget_charge(b, held());
        b->ctxs.push_back(c);
        c->budget = b;
    }
    return NMFG_OK;
}

int nmfg_device_memory(int32_t device, int64_t *free_bytes, int64_t *total_bytes) {
    if (!free_bytes || !total_bytes) return fail(NMFG_EINVAL, "NULL argument");
    int prev = 0;
    HIPCHK(hipGetDevice(&prev));
    HIPCHK(hipSetDevice(device));
    size_t f = 0, t = 0;
    const hipError_t e = hipMemGetInfo(&f, &t);
    (void)hipSetDevice(prev);
    if (e != hipSuccess) return fail(NMFG_EDEVICE, std::string("hipMemGetInfo: ") + hipGetErrorString(e));
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
    return NMFG_OK;
}

int nmfg_budget_query(nmfg_budget *b, int64_t *cap, int64_t *used, int64_t *peak) {
    if (!b) return fail(NMFG_EINVAL, "NULL budget");
    std::lock_guard<std::mutex> lk(b->mu);
    if (cap) *cap = b->cap;
    if (used) *used = b->used;
    if (peak) *peak = b->peak;
    return NMFG_OK;
}

int nmfg_matrix_create(nmfg_ctx *ctx, int64_t n, int64_t p, int32_t precision, nmfg_matrix **out) {
    if (!ctx || !out) return fail(NMFG_EINVAL, "NULL argument");
    if (n < 1 || p < 1) return fail(NMFG_EINVAL, "V must have at least one row and one column");
    if (n > (int64_t)1 << 30 || p > (int64_t)1 << 30) return fail(NMFG_EINVAL, "V too large");
    if (!valid_form(precision)) return fail(NMFG_EINVAL, "unknown precision code");
    Guard g(ctx);
    nmfg_matrix *m = new nmfg_matrix();
    m->ctx = ctx;
    m->n = n;
    m->p = p;
    m->ldv = (p + LDV_ALIGN - 1) / LDV_ALIGN * LDV_ALIGN;
    m->form = precision;
    m->precision = storage_of(precision);
    m->vplane = m->precision == VP_B16X2 ? n * m->ldv : 0;
    size_t vc_off = 0;
    const size_t bytes = v_block_bytes(precision, n, m->ldv, &vc_off);
    int rc = ctx_alloc(ctx, bytes, &m->V);
    if (rc) {
        delete m;
        return rc;
    }
    if (precision == NMFG_BF16X2) {
        m->nc = b16_rows(n);
        m->Vc = reinterpret_cast<uint16_t *>(static_cast<char *>(m->V) + vc_off);
    }
    hipError_t e = hipMemsetAsync(m->V, 0, bytes, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        ctx_release(ctx, m->V);
        delete m;
        return fail(NMFG_EDEVICE, std::string("hipMemset(V): ") + hipGetErrorString(e));
    }
    *out = m;
    return NMFG_OK;
}

int nmfg_matrix_upload(nmfg_matrix *m, const double *V, int64_t ldv) {
    if (!m || !V) return fail(NMFG_EINVAL, "NULL argument");
    if (ldv < m->p) return fail(NMFG_EINVAL, "ldv < p");
    Guard g(m->ctx);
    hipStream_t st = m->ctx->stream;
    if (m->precision == VP_F64) {
        HIPCHK(hipMemcpy2DAsync(m->V, sizeof(double) * m->ldv, V, sizeof(double) * ldv,
                                sizeof(double) * m->p, m->n, hipMemcpyDefault, st));
        HIPCHK(hipStreamSynchronize(st));
        return NMFG_OK;
    }
    // Other forms: FP64 row chunks through a device staging buffer, converted
    // on the device.
    const int64_t chunk_bytes = (int64_t)256 << 20;
    int64_t rows_per = std::max<int64_t>(1, chunk_bytes / (int64_t)(sizeof(double) * m->p));
    // BF16 form: chunks of whole 8-row blocks for the column pass's copy
    if (m->Vc) rows_per = std::max<int64_t>(8, rows_per / 8 * 8);
    const int64_t rows_alloc = std::min<int64_t>(rows_per, m->n);
    double *stage = nullptr;
    int rc = ctx_alloc(m->ctx, sizeof(double) * (size_t)rows_alloc * m->p, (void **)&stage);
    if (rc) return rc;
    hipError_t e = hipSuccess;
    const int64_t esz = m->precision == VP_F32 ? 4 : 2;
    for (int64_t r0 = 0; r0 < m->n && e == hipSuccess; r0 += rows_alloc) {
        const int64_t rows = std::min<int64_t>(rows_alloc, m->n - r0);
        e = hipMemcpy2DAsync(stage, sizeof(double) * m->p, V + r0 * ldv, sizeof(double) * ldv,
                             sizeof(double) * m->p, rows, hipMemcpyDefault, st);
        if (e == hipSuccess)
            e = launch_convert_v(m->precision, stage, m->p, rows, m->p,
                                 static_cast<char *>(m->V) + (size_t)(r0 * m->ldv * esz), m->ldv,
                                 m->vplane, st);
        if (e == hipSuccess && m->Vc)
            e = launch_convert_vc(stage, m->p, rows, r0, m->p, m->Vc, m->ldv, m->nc * m->ldv, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    ctx_release(m->ctx, stage);
    if (e != hipSuccess) return fail(NMFG_EDEVICE, std::string("upload: ") + hipGetErrorString(e));
    return NMFG_OK;
}

int nmfg_matrix_destroy(nmfg_matrix *m) {
    if (!m) return NMFG_OK;
    Guard g(m->ctx);
    (void)hipStreamSynchronize(m->ctx->stream);
    for (nmfg_solve *&ps : m->prod)
        if (ps) {
            solve_free(ps);
            delete ps;
            ps = nullptr;
        }
    ctx_release(m->ctx, m->mul_buf);
    ctx_release(m->ctx, m->V);
    delete m;
    return NMFG_OK;
}

int nmfg_solve_create(nmfg_matrix *m, int32_t r, const nmfg_params *prm, nmfg_solve **out) {
    if (!m || !out) return fail(NMFG_EINVAL, "NULL argument");
    int rc = validate_params(prm);
    if (rc) return rc;
    Guard g(m->ctx);
    return solve_create_kind(m, r, prm, kind_of(prm->algo), 1, out);
}

int nmfg_solve_create_batched(nmfg_matrix *m, int32_t r, const nmfg_params *prm, int32_t batch,
                              nmfg_solve **out) {
    if (!m || !out) return fail(NMFG_EINVAL, "NULL argument");
    int rc = validate_params(prm);
    if (rc) return rc;
    Guard g(m->ctx);
    return solve_create_kind(m, r, prm, kind_of(prm->algo), batch, out);
}

int32_t nmfg_solve_batch(const nmfg_solve *s) { return s ? s->batch : 0; }

int nmfg_consensus(nmfg_ctx *ctx, const int32_t *labels, int32_t runs, int64_t p, double *C) {
    if (!ctx || !labels || !C) return fail(NMFG_EINVAL, "NULL argument");
    if (runs < 1 || p < 1) return fail(NMFG_EINVAL, "runs and p must be >= 1");
    if (p > (int64_t)1 << 20) return fail(NMFG_EINVAL, "p too large for a dense consensus matrix");
    Guard g(ctx);
    hipStream_t st = ctx->stream;
    int32_t *dl = nullptr;
    double *dC = nullptr;
    const size_t lb = sizeof(int32_t) * (size_t)runs * p, cb = sizeof(double) * (size_t)p * p;
    int rc = ctx_alloc(ctx, lb, (void **)&dl);
    if (!rc) rc = ctx_alloc(ctx, cb, (void **)&dC);
    if (rc) {
        ctx_release(ctx, dl);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(dl, labels, lb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_consensus(dl, runs, (int)p, dC, st);
    if (e == hipSuccess) e = hipMemcpyAsync(C, dC, cb, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    ctx_release(ctx, dl);
    ctx_release(ctx, dC);
    if (e != hipSuccess) return fail(NMFG_EDEVICE, std::string("consensus: ") + hipGetErrorString(e));
    return NMFG_OK;
}

int nmfg_solve_set_factors(nmfg_solve *s, const double *W0, const double *H0) {
    if (!s || !W0 || !H0) return fail(NMFG_EINVAL, "NULL argument");
    Guard g(s->m->ctx);
    int rc = pack_factors(s, W0, H0, s->W0, s->H0);
    if (rc) return rc;
    s->have_factors = true;
    return NMFG_OK;
}

int nmfg_solve_set_weight(nmfg_solve *s, const double *weight, int64_t ldw) {
    if (!s || !weight) return fail(NMFG_EINVAL, "NULL argument");
    if (s->kind != K_LSNMF) return fail(NMFG_EINVAL, "the solve's method is not lsNMF");
    nmfg_matrix *m = s->m;
    if (ldw < m->p) return fail(NMFG_EINVAL, "ldw < p");
    Guard g(m->ctx);
    hipStream_t st = m->ctx->stream;
    // one byte per entry when every weight is 0 or 1, else FP64; zero padded to ldv
    bool binary = true;
    for (int64_t i = 0; i < m->n && binary; ++i)
        for (int64_t j = 0; j < m->p; ++j) {
            const double x = weight[i * ldw + j];
            if (x != 0.0 && x != 1.0) {
                binary = false;
                break;
            }
        }
    const int wf = binary ? WF_U8 : WF_F64;
    const size_t esz = binary ? 1 : sizeof(double);
    const size_t bytes = esz * (size_t)m->n * m->ldv;
    if (s->wgt && s->wf != wf) {
        for (auto it = s->allocs.begin(); it != s->allocs.end(); ++it)
            if (*it == s->wgt) {
                s->allocs.erase(it);
                break;
            }
        ctx_release(m->ctx, s->wgt);
        s->wgt = nullptr;
    }
    if (!s->wgt) {
        int rc = dmalloc(s, &s->wgt, bytes);
        if (rc) return rc;
    }
    s->wf = wf;
    std::vector<uint8_t> host(bytes, 0);
    for (int64_t i = 0; i < m->n; ++i)
        for (int64_t j = 0; j < m->p; ++j) {
            const double x = weight[i * ldw + j];
            if (binary)
                host[(size_t)i * m->ldv + j] = x != 0.0 ? 1 : 0;
            else
                std::memcpy(&host[((size_t)i * m->ldv + j) * sizeof(double)], &x, sizeof(double));
        }
    HIPCHK(hipMemcpyAsync(s->wgt, host.data(), bytes, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    // The captured graphs (the fixed-iteration solve and the objective-stop
    // chunk) hold the previous weight's address and form.
    if (s->graph) {
        (void)hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
    }
    if (s->chunk_graph) {
        (void)hipGraphExecDestroy(s->chunk_graph);
        s->chunk_graph = nullptr;
    }
    s->chunk_graph_failed = false;
    return NMFG_OK;
}

int nmfg_solve_set_offset(nmfg_solve *s, const double *offset) {
    if (!s || !offset) return fail(NMFG_EINVAL, "NULL argument");
    if (s->kind != K_OFFSET) return fail(NMFG_EINVAL, "the solve's method is not offset");
    Guard g(s->m->ctx);
    hipStream_t st = s->m->ctx->stream;
    HIPCHK(hipMemcpyAsync(s->off0, offset, sizeof(double) * (size_t)s->m->n * s->batch,
                          hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return NMFG_OK;
}

int nmfg_solve_get_offset(nmfg_solve *s, double *offset) {
    if (!s || !offset) return fail(NMFG_EINVAL, "NULL argument");
    if (s->kind != K_OFFSET) return fail(NMFG_EINVAL, "the solve's method is not offset");
    Guard g(s->m->ctx);
    hipStream_t st = s->m->ctx->stream;
    HIPCHK(hipMemcpyAsync(offset, s->off, sizeof(double) * (size_t)s->m->n * s->batch,
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return NMFG_OK;
}

int nmfg_solve_launch(nmfg_solve *s) {
    if (!s) return fail(NMFG_EINVAL, "NULL argument");
    if (!s->have_factors) return fail(NMFG_EINVAL, "set W0/H0 first");
    if (s->kind == K_LSNMF && !s->wgt)
        return fail(NMFG_EINVAL, "method='lsnmf' requires a `weight` matrix of the same shape as V");
    Guard g(s->m->ctx);
    hipStream_t st = s->m->ctx->stream;
    if (s->prm.stop != NMFG_STOP_MAX_ITER) return run_with_objective(s, st);
    s->n_iter = s->prm.max_iter;
    s->devs.clear();
    if (env_int("NMFG_NO_GRAPH", 0)) return enqueue_fixed(s, st);
    if (!s->graph) {
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_fixed(s, st);
        hipError_t e = hipStreamEndCapture(st, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return fail(NMFG_EDEVICE, std::string("capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&s->graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            s->graph = nullptr;
            return fail(NMFG_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
    }
    HIPCHK(hipGraphLaunch(s->graph, st));
    return NMFG_OK;
}

int nmfg_solve_fetch(nmfg_solve *s, double *W, double *H, int64_t *n_iter, double *dev,
                     int64_t dev_cap, int64_t *n_dev) {
    if (!s) return fail(NMFG_EINVAL, "NULL argument");
    Guard g(s->m->ctx);
    HIPCHK(hipStreamSynchronize(s->m->ctx->stream));
    int rc = unpack_factors(s, s->W, s->H, W, H);
    if (rc) return rc;
    if (n_iter) *n_iter = s->n_iter;
    if (n_dev) *n_dev = (int64_t)s->devs.size();
    if (dev)
        for (int64_t i = 0; i < (int64_t)s->devs.size() && i < dev_cap; ++i) dev[i] = s->devs[i];
    return NMFG_OK;
}

namespace {

// What time_pass_graph puts before each pass: nothing, or a read of an
// eviction buffer (ev.n16 16-byte words) -- or that read alone (pass false).
struct EvictPlan {
    const void *buf = nullptr;
    int64_t n16 = 0;
    unsigned *out = nullptr;
    bool pass = true;
};

// Time `reps` back-to-back launches of one pass captured in a graph (no host
// work or events between the launches, as inside a solve's graph).
int time_pass_graph(nmfg_solve *s, hipStream_t st, bool col, int64_t reps, double *ms,
                    const EvictPlan &ev = EvictPlan{}) {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = NMFG_OK;
    for (int64_t i = 0; i < reps && !rc; ++i) {
        if (ev.buf && launch_evict(ev.buf, ev.n16, ev.out, st) != hipSuccess) rc = NMFG_EDEVICE;
        if (!rc && ev.pass) rc = col ? launch_col(s, st, 0, 0) : launch_row(s, st, 0);
    }
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc ? rc : fail(NMFG_EDEVICE, std::string("capture: ") + hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(NMFG_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipGraphLaunch(exec, st);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    if (e == hipSuccess) e = hipGraphLaunch(exec, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float x = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&x, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipGraphExecDestroy(exec);
    if (e != hipSuccess) return fail(NMFG_EDEVICE, std::string("profile: ") + hipGetErrorString(e));
    *ms = x;
    return NMFG_OK;
}

}  // namespace

int nmfg_solve_profile(nmfg_solve *s, double *col_ms, int64_t *col_n, double *row_ms,
                       int64_t *row_n) {
    if (!s) return fail(NMFG_EINVAL, "NULL argument");
    if (!s->have_factors) return fail(NMFG_EINVAL, "set W0/H0 first");
    Guard g(s->m->ctx);
    hipStream_t st = s->m->ctx->stream;
    const int64_t reps = std::max<int64_t>(1, s->prm.max_iter);
    // One real iteration first so that every statistic the passes read is set.
    int rc = enqueue_start(s, st);
    if (!rc) rc = enqueue_iteration(s, st, 1);
    double cm = 0.0, rm = 0.0;
    if (!rc) rc = time_pass_graph(s, st, true, reps, &cm);
    if (!rc) rc = launch_red(s, st, s->ntiles_col);
    if (!rc) rc = time_pass_graph(s, st, false, reps, &rm);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(st));
    s->n_iter = 0;
    if (col_ms) *col_ms = cm;
    if (row_ms) *row_ms = rm;
    if (col_n) *col_n = reps;
    if (row_n) *row_n = reps;
    return NMFG_OK;
}

int nmfg_solve_profile_evicted(nmfg_solve *s, int64_t evict_bytes, double *col_ms, int64_t *col_n,
                               double *row_ms, int64_t *row_n) {
    if (!s) return fail(NMFG_EINVAL, "NULL argument");
    if (!s->have_factors) return fail(NMFG_EINVAL, "set W0/H0 first");
    if (evict_bytes < ((int64_t)1 << 20) || evict_bytes > ((int64_t)64 << 30))
        return fail(NMFG_EINVAL, "evict_bytes must be in [1 MiB, 64 GiB]");
    Guard g(s->m->ctx);
    nmfg_ctx *c = s->m->ctx;
    hipStream_t st = c->stream;
    const int64_t reps = std::max<int64_t>(1, s->prm.max_iter);
    const int64_t n16 = evict_bytes / 16;
    const size_t out_bytes = sizeof(unsigned) * EVICT_BLOCKS * 256;
    void *buf = nullptr;
    int rc = ctx_alloc(c, (size_t)n16 * 16 + out_bytes, &buf);
    if (rc) return rc;
    EvictPlan ev;
    ev.buf = buf;
    ev.n16 = n16;
    ev.out = reinterpret_cast<unsigned *>(static_cast<char *>(buf) + (size_t)n16 * 16);
    EvictPlan alone = ev;
    alone.pass = false;
    hipError_t e = hipMemsetAsync(buf, 0, (size_t)n16 * 16, st);
    double cm = 0.0, rm = 0.0, em = 0.0;
    if (e != hipSuccess) rc = fail(NMFG_EDEVICE, std::string("hipMemset: ") + hipGetErrorString(e));
    if (!rc) rc = enqueue_start(s, st);
    if (!rc) rc = enqueue_iteration(s, st, 1);
    if (!rc) rc = time_pass_graph(s, st, true, reps, &cm, ev);
    if (!rc) rc = launch_red(s, st, s->ntiles_col);
    if (!rc) rc = time_pass_graph(s, st, false, reps, &rm, ev);
    if (!rc) rc = time_pass_graph(s, st, false, reps, &em, alone);
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = fail(NMFG_EDEVICE, "profile sync");
    ctx_release(c, buf);
    if (rc) return rc;
    s->n_iter = 0;
    if (col_ms) *col_ms = std::max(0.0, cm - em);
    if (row_ms) *row_ms = std::max(0.0, rm - em);
    if (col_n) *col_n = reps;
    if (row_n) *row_n = reps;
    return NMFG_OK;
}

namespace {

// C = V B (trans 0; B p x k, C n x k) or V^T B (trans 1; B n x k, C p x k),
// row-major device buffers, enqueued on the matrix's stream.  Each chunk of
// at most 32 columns (balanced widths) is written into a K_PROD solve's
// factor panel, one pass forms the product panel, which is scattered into C.
int mul_device(nmfg_matrix *m, int trans, double *dB, int k, double *dC) {
    hipStream_t st = m->ctx->stream;
    const int64_t n = m->n, p = m->p, ldv = m->ldv;
    const int64_t rin = trans ? n : p, rout = trans ? p : n;
    // Chunks of 32 columns, then the remainder: a pass costs by the 16-wide
    // MFMA tiles of its template rank, two for any chunk of 17..32 columns,
    // so 34 columns as 32 + 2 take three tiles over two passes of V where
    // 17 + 17 take four.
    const int nch = (k + 31) / 32;
    int c0 = 0;
    for (int c = 0; c < nch; ++c) {
        const int kc = std::min(32, k - c0);
        nmfg_solve *&s = m->prod[kc];
        int rc = NMFG_OK;
        if (!s) {
            nmfg_params prm;
            nmfg_params_init(&prm);
            prm.max_iter = 0;
            if ((rc = solve_create_kind(m, kc, &prm, K_PROD, 1, &s))) return rc;
        }
        // trans 0: B's chunk as the H panel (R x ldv), row pass into W (n x R);
        // trans 1: B's chunk as the W panel, column pass into H
        PanelArgs in{dB, rin, k, c0, kc, trans ? s->W : s->H, s->R, trans ? 0 : ldv, trans ? 0 : 1, 1};
        HIPCHK(launch_panel(in, st));
        if ((rc = trans ? launch_col(s, st, 1, 0) : launch_row(s, st, 0))) return rc;
        PanelArgs out{dC, rout, k, c0, kc, trans ? s->H : s->W, s->R, trans ? ldv : 0, trans ? 1 : 0, 0};
        HIPCHK(launch_panel(out, st));
        c0 += kc;
    }
    return NMFG_OK;
}

// Device buffer of at least `bytes` kept with the matrix (grown on demand).
int matrix_scratch(nmfg_matrix *m, size_t bytes, void **out) {
    if (bytes > m->mul_cap) {
        ctx_release(m->ctx, m->mul_buf);
        m->mul_buf = nullptr;
        m->mul_cap = 0;
        int rc = ctx_alloc(m->ctx, bytes, (void **)&m->mul_buf);
        if (rc) return rc;
        m->mul_cap = bytes;
    }
    *out = m->mul_buf;
    return NMFG_OK;
}

}  // namespace

int nmfg_matrix_mul(nmfg_matrix *m, int32_t trans, const double *B, int32_t k, double *C) {
    if (!m || !B || !C) return fail(NMFG_EINVAL, "NULL argument");
    if (k < 1) return fail(NMFG_EINVAL, "k must be >= 1");
    if (trans != 0 && trans != 1) return fail(NMFG_EINVAL, "trans must be 0 or 1");
    Guard g(m->ctx);
    hipStream_t st = m->ctx->stream;
    const int64_t rin = trans ? m->n : m->p, rout = trans ? m->p : m->n;
    void *buf = nullptr;
    int rc = matrix_scratch(m, sizeof(double) * (size_t)(rin + rout) * k, &buf);
    if (rc) return rc;
    double *dB = static_cast<double *>(buf), *dC = dB + (size_t)rin * k;
    HIPCHK(hipMemcpyAsync(dB, B, sizeof(double) * (size_t)rin * k, hipMemcpyHostToDevice, st));
    if ((rc = mul_device(m, trans, dB, k, dC))) return rc;
    HIPCHK(hipMemcpyAsync(C, dC, sizeof(double) * (size_t)rout * k, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return NMFG_OK;
}

int nmfg_matrix_range_finder(nmfg_matrix *m, int32_t flip, const double *Q0, int32_t width,
                             int32_t n_iter, double *Y) {
    if (!m || !Q0 || !Y) return fail(NMFG_EINVAL, "NULL argument");
    if (flip != 0 && flip != 1) return fail(NMFG_EINVAL, "flip must be 0 or 1");
    if (width < 1 || n_iter < 0) return fail(NMFG_EINVAL, "width must be >= 1 and n_iter >= 0");
    // A = V^T (flip) or V, a x b; Q0 is b x width
    const int64_t a = flip ? m->p : m->n, b = flip ? m->n : m->p;
    if (a < width || b < width)
        return fail(NMFG_EINVAL, "both dimensions of V must be >= width for the LU steps");
    if (std::max(a, b) >= (int64_t)INT_MAX) return fail(NMFG_EINVAL, "V too large for the LU steps");
    Guard g(m->ctx);
    hipStream_t st = m->ctx->stream;
    const size_t qa = align256(sizeof(double) * (size_t)a * width);
    const size_t qb = align256(sizeof(double) * (size_t)b * width);
    void *buf = nullptr;
    int rc = matrix_scratch(m, qa + qb + lu_work_bytes(std::max(a, b), width), &buf);
    if (rc) return rc;
    char *base = static_cast<char *>(buf);
    double *dA = reinterpret_cast<double *>(base);       // a x width: A Q
    double *dB = reinterpret_cast<double *>(base + qa);  // b x width: A^T Q
    const LuWork ws = lu_work_at(base + qa + qb, std::max(a, b), width);
    HIPCHK(hipMemcpyAsync(dB, Q0, sizeof(double) * (size_t)b * width, hipMemcpyHostToDevice, st));
    // A Q is V^T Q (trans 1) when flipped, V Q (trans 0) otherwise
    const int fwd = flip ? 1 : 0, back = flip ? 0 : 1;
    for (int it = 0; it < n_iter; ++it) {
        if ((rc = mul_device(m, fwd, dB, width, dA))) return rc;
        HIPCHK(launch_lu_pl(dA, a, width, ws, st));
        if ((rc = mul_device(m, back, dA, width, dB))) return rc;
        HIPCHK(launch_lu_pl(dB, b, width, ws, st));
    }
    if ((rc = mul_device(m, fwd, dB, width, dA))) return rc;
    HIPCHK(hipMemcpyAsync(Y, dA, sizeof(double) * (size_t)a * width, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return NMFG_OK;
}

int nmfg_lu_permute_l(nmfg_ctx *ctx, const double *X, int64_t m, int32_t w, double *PL) {
    if (!ctx || !X || !PL) return fail(NMFG_EINVAL, "NULL argument");
    if (w < 1 || m < w) return fail(NMFG_EINVAL, "need m >= w >= 1");
    if (m >= (int64_t)INT_MAX) return fail(NMFG_EINVAL, "m too large");
    Guard g(ctx);
    hipStream_t st = ctx->stream;
    const size_t xb = align256(sizeof(double) * (size_t)m * w);
    void *buf = nullptr;
    int rc = ctx_alloc(ctx, xb + lu_work_bytes(m, w), &buf);
    if (rc) return rc;
    double *dX = static_cast<double *>(buf);
    const LuWork ws = lu_work_at(static_cast<char *>(buf) + xb, m, w);
    hipError_t e = hipMemcpyAsync(dX, X, sizeof(double) * (size_t)m * w, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_lu_pl(dX, m, w, ws, st);
    if (e == hipSuccess) e = hipMemcpyAsync(PL, dX, sizeof(double) * (size_t)m * w, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    ctx_release(ctx, buf);
    if (e != hipSuccess) return fail(NMFG_EDEVICE, std::string("lu: ") + hipGetErrorString(e));
    return NMFG_OK;
}

int nmfg_solve_destroy(nmfg_solve *s) {
    if (!s) return NMFG_OK;
    Guard g(s->m->ctx);
    (void)hipStreamSynchronize(s->m->ctx->stream);
    solve_free(s);
    delete s;
    return NMFG_OK;
}

// ---------------------------------------------------------------------------
// One-call entry points (host buffers in, host buffers out).
// ---------------------------------------------------------------------------
namespace {

struct Scoped {
    nmfg_matrix *m = nullptr;
    nmfg_solve *s = nullptr;
    ~Scoped() {
        if (s) nmfg_solve_destroy(s);
        if (m) nmfg_matrix_destroy(m);
    }
};

int check_host_shapes(const double *V, int64_t n, int64_t p, const void *W, const void *H,
                      int32_t r) {
    if (!V || !W || !H) return fail(NMFG_EINVAL, "NULL matrix argument");
    if (n < 1 || p < 1) return fail(NMFG_EINVAL, "V must be non-empty");
    if (r < 1) return fail(NMFG_EINVAL, "rank must be >= 1");
    return NMFG_OK;
}

// One half-update through the pass kernels on fresh statistics.
int single_step(nmfg_ctx *ctx, int32_t algo, bool h_side, const double *V, int64_t n, int64_t p,
                const double *W, const double *H, int32_t r, double eps, double *out) {
    int rc = check_host_shapes(V, n, p, W, H, r);
    if (rc) return rc;
    if (!ctx || !out) return fail(NMFG_EINVAL, "NULL argument");
    nmfg_params prm;
    nmfg_params_init(&prm);
    prm.algo = algo;
    prm.eps = eps;
    prm.max_iter = 0;
    prm.rescale = 0;
    Guard g(ctx);
    Scoped sc;
    if ((rc = nmfg_matrix_create(ctx, n, p, NMFG_FP64, &sc.m))) return rc;
    if ((rc = nmfg_matrix_upload(sc.m, V, p))) return rc;
    if ((rc = nmfg_solve_create(sc.m, r, &prm, &sc.s))) return rc;
    if ((rc = nmfg_solve_set_factors(sc.s, W, H))) return rc;
    hipStream_t st = ctx->stream;
    if ((rc = enqueue_start(sc.s, st))) return rc;
    if (h_side) {
        if ((rc = launch_col(sc.s, st, 1, 0))) return rc;
    } else {
        if ((rc = launch_hstats(sc.s, st)) || (rc = launch_row(sc.s, st, 0))) return rc;
    }
    return h_side ? unpack_factors(sc.s, sc.s->W, sc.s->H, nullptr, out)
                  : unpack_factors(sc.s, sc.s->W, sc.s->H, out, nullptr);
}

}  // namespace

int nmfg_euclidean_update_h(nmfg_ctx *ctx, const double *V, int64_t n, int64_t p, const double *W,
                            const double *H, int32_t r, double eps, double *H_out) {
    return single_step(ctx, NMFG_LEE, true, V, n, p, W, H, r, eps, H_out);
}

int nmfg_euclidean_update_w(nmfg_ctx *ctx, const double *V, int64_t n, int64_t p, const double *W,
                            const double *H, int32_t r, double eps, double *W_out) {
    return single_step(ctx, NMFG_LEE, false, V, n, p, W, H, r, eps, W_out);
}

int nmfg_divergence_update_h(nmfg_ctx *ctx, const double *V, int64_t n, int64_t p,
                             const double *W, const double *H, int32_t r, double *H_out) {
    return single_step(ctx, NMFG_BRUNET, true, V, n, p, W, H, r, DBL_EPSILON, H_out);
}

int nmfg_divergence_update_w(nmfg_ctx *ctx, const double *V, int64_t n, int64_t p,
                             const double *W, const double *H, int32_t r, double *W_out) {
    return single_step(ctx, NMFG_BRUNET, false, V, n, p, W, H, r, DBL_EPSILON, W_out);
}

}  // extern "C"

namespace {

// Kim-Park sparse NMF, snmf/r and snmf/l (lib.rs:1198-1334; dispatch
// 2117-2146): alternating FCNNLS on stacked systems.  With A = V (snmf/r) or
// V^T (snmf/l), internal factors Wi (m x r), Hi (r x n):
//   H step: [Wi; sqrt(beta) 1^T] K = [A; 0]   -> normal equations
//           Wi^T Wi + beta 1 1^T,  Wi^T A
//   W step: [Hi^T; eta I] K = [A^T; 0]        -> Hi Hi^T + eta^2 I,  Hi A^T
// The products with V (W^T V, V H^T) and the objective run on the GPU
// (product-mode passes, k_obj); the r x r Grams and FCNNLS are FP64 on the
// host.  One deviance 0.5 ||V - W H||^2 per iteration, no early stop.
struct SnmfCtx {
    nmfg_solve *s;
    hipStream_t st;
    int64_t n, p;
    int r;
};

// out (n x r, row-major) = V Hp^T for Hp (r x p).
int snmf_vht(SnmfCtx &c, const double *Hp, double *out) {
    int rc = pack_factors(c.s, nullptr, Hp, nullptr, c.s->H);
    if (!rc) rc = launch_row(c.s, c.st, 0);
    if (!rc) rc = unpack_factors(c.s, c.s->W, c.s->H, out, nullptr);
    return rc;
}

// out (r x p, row-major) = Wp^T V for Wp (n x r).
int snmf_wtv(SnmfCtx &c, const double *Wp, double *out) {
    int rc = pack_factors(c.s, Wp, nullptr, c.s->W, nullptr);
    if (!rc) rc = launch_col(c.s, c.st, 1, 0);
    if (!rc) rc = unpack_factors(c.s, c.s->W, c.s->H, nullptr, out);
    return rc;
}

// 0.5 ||V - W H||^2 on the GPU (per-tile partials summed here in order).
int snmf_objective(SnmfCtx &c, const double *W, const double *H, double *f) {
    nmfg_solve *s = c.s;
    int rc = pack_factors(s, W, H, s->W, s->H);
    if (rc) return rc;
    ObjArgs o{};
    nmfg_matrix *m = s->m;
    o.V = m->V;
    o.vplane = m->vplane;
    o.ldv = m->ldv;
    o.n = (int)m->n;
    o.p = (int)m->p;
    o.r = s->r;
    o.ntiles = s->ntiles_obj;
    o.osplit = s->osplit;
    o.W = s->W;
    o.H = s->H;
    o.wsc = s->wsc;
    o.unit_sc = 1;  // W as given
    o.hst = s->hst;
    o.part = s->obj_part;
    o.mfma = s->obj_mfma ? 1 : 0;
    HIPCHK(s->K->obj(o, c.st));
    std::vector<double> part((size_t)s->ntiles_obj * s->osplit);
    HIPCHK(hipMemcpyAsync(part.data(), s->obj_part, sizeof(double) * part.size(),
                          hipMemcpyDeviceToHost, c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    double t = 0.0;
    for (double x : part) t += x;
    *f = 0.5 * t;
    return NMFG_OK;
}

// G (r x r) = X^T X of the rows of X (rows x r), forward over the rows.
void gram_rows(const double *X, int64_t rows, int r, double *G) {
    for (int a = 0; a < r; ++a)
        for (int b = 0; b < r; ++b) {
            double s = 0.0;
            for (int64_t i = 0; i < rows; ++i) s += X[i * r + a] * X[i * r + b];
            G[a * r + b] = s;
        }
}

int run_snmf(nmfg_ctx *ctx, const nmfg_params *prm, const double *V, int64_t n, int64_t p,
             int64_t ldv, double *W, double *H, int32_t r, int64_t *n_iter, double *dev,
             int64_t dev_cap, int64_t *n_dev) {
    if (prm->stop != NMFG_STOP_MAX_ITER && prm->stop != NMFG_STOP_STATIONARY &&
        prm->stop != NMFG_STOP_REL_CHANGE)
        return fail(NMFG_EINVAL, "unknown stopping criterion");
    if (prm->max_iter < 0) return fail(NMFG_EINVAL, "max_iter must be >= 0");
    if (!valid_form(prm->precision)) return fail(NMFG_EINVAL, "unknown precision code");
    if (r > 32) return fail(NMFG_EUNSUPPORTED, "snmf supports ranks up to 32");
    const bool left = prm->algo == NMFG_SNMF_L;
    Guard g(ctx);
    Scoped sc;
    int rc;
    if ((rc = nmfg_matrix_create(ctx, n, p, prm->precision, &sc.m))) return rc;
    if ((rc = nmfg_matrix_upload(sc.m, V, ldv))) return rc;
    nmfg_params pp;
    nmfg_params_init(&pp);
    pp.max_iter = 0;
    if ((rc = solve_create_kind(sc.m, r, &pp, K_PROD, 1, &sc.s))) return rc;
    SnmfCtx c{sc.s, ctx->stream, n, p, r};
    // eta < 0 means max(A), folded from 0.0 (lib.rs:1228-1232)
    double eta = left ? prm->smoothness_h : prm->smoothness_w;
    if (eta < 0.0) {
        eta = 0.0;
        for (int64_t i = 0; i < n; ++i)
            for (int64_t j = 0; j < p; ++j) eta = std::fmax(eta, V[i * ldv + j]);
    }
    const double sb = std::sqrt(std::fmax(left ? prm->sparsity_w : prm->sparsity_h, 1e-30));
    const double eps = prm->eps;
    const int64_t mi = left ? p : n, ni = left ? n : p;  // internal A is mi x ni
    std::vector<double> Wi((size_t)mi * r), Hi((size_t)r * ni);
    if (!left) {
        std::memcpy(Wi.data(), W, sizeof(double) * Wi.size());
        std::memcpy(Hi.data(), H, sizeof(double) * Hi.size());
    } else {
        for (int64_t j = 0; j < p; ++j)
            for (int k = 0; k < r; ++k) Wi[(size_t)j * r + k] = H[(size_t)k * p + j];
        for (int k = 0; k < r; ++k)
            for (int64_t i = 0; i < n; ++i) Hi[(size_t)k * n + i] = W[(size_t)i * r + k];
    }
    for (int k = 0; k < r; ++k) {  // col_l2_normalize (lib.rs:1188-1197)
        double s = 0.0;
        for (int64_t i = 0; i < mi; ++i) s += Wi[(size_t)i * r + k] * Wi[(size_t)i * r + k];
        const double nrm = std::sqrt(s);
        if (nrm > 0.0)
            for (int64_t i = 0; i < mi; ++i) Wi[(size_t)i * r + k] = Wi[(size_t)i * r + k] / nrm;
    }
    std::vector<double> ctc((size_t)r * r), rhs((size_t)std::max(mi, ni) * r), prod, sol,
        HiT((size_t)ni * r), Wu((size_t)n * r), Hu((size_t)r * p);
    prod.resize((size_t)std::max(n, p) * r);
    sol.resize((size_t)std::max(mi, ni) * r);
    int64_t nd = 0;
    for (int64_t it = 0; it < prm->max_iter; ++it) {
        // H step: rhs row j = (Wi^T A)[:, j]
        gram_rows(Wi.data(), mi, r, ctc.data());
        for (int a = 0; a < r * r; ++a) ctc[a] += sb * sb;
        if (!left) {  // Wi^T V (r x p) -> rows j
            if ((rc = snmf_wtv(c, Wi.data(), prod.data()))) return rc;
            for (int64_t j = 0; j < p; ++j)
                for (int k = 0; k < r; ++k) rhs[(size_t)j * r + k] = prod[(size_t)k * p + j];
        } else {  // (V Wi)^T with Wi = H_user^T: V H_user^T (n x r) is already rows j
            for (int k = 0; k < r; ++k)
                for (int64_t j = 0; j < p; ++j) Hu[(size_t)k * p + j] = Wi[(size_t)j * r + k];
            if ((rc = snmf_vht(c, Hu.data(), rhs.data()))) return rc;
        }
        if (fcnnls_normal(ctc.data(), rhs.data(), r, ni, eps, sol.data()))
            return fail(NMFG_EINVAL, "fcnnls: rank out of range");
        for (int64_t j = 0; j < ni; ++j)
            for (int k = 0; k < r; ++k) Hi[(size_t)k * ni + j] = sol[(size_t)j * r + k];
        // W step: rhs row i = (Hi A^T)[:, i]
        for (int64_t j = 0; j < ni; ++j)
            for (int k = 0; k < r; ++k) HiT[(size_t)j * r + k] = Hi[(size_t)k * ni + j];
        gram_rows(HiT.data(), ni, r, ctc.data());
        for (int a = 0; a < r; ++a) ctc[a * r + a] += eta * eta;
        if (!left) {  // Hi V^T = (V Hi^T)^T: V Hi^T (n x r) is rows i
            if ((rc = snmf_vht(c, Hi.data(), rhs.data()))) return rc;
        } else {  // Hi V with Hi = W_user^T: W_user^T V (r x p) -> rows i
            if ((rc = snmf_wtv(c, HiT.data(), prod.data()))) return rc;
            for (int64_t i = 0; i < p; ++i)
                for (int k = 0; k < r; ++k) rhs[(size_t)i * r + k] = prod[(size_t)k * p + i];
        }
        if (fcnnls_normal(ctc.data(), rhs.data(), r, mi, eps, sol.data()))
            return fail(NMFG_EINVAL, "fcnnls: rank out of range");
        std::memcpy(Wi.data(), sol.data(), sizeof(double) * Wi.size());
        // user factors and the deviance
        if (!left) {
            std::memcpy(Wu.data(), Wi.data(), sizeof(double) * Wu.size());
            std::memcpy(Hu.data(), Hi.data(), sizeof(double) * Hu.size());
        } else {
            std::memcpy(Wu.data(), HiT.data(), sizeof(double) * Wu.size());
            for (int k = 0; k < r; ++k)
                for (int64_t j = 0; j < p; ++j) Hu[(size_t)k * p + j] = Wi[(size_t)j * r + k];
        }
        double f = 0.0;
        if ((rc = snmf_objective(c, Wu.data(), Hu.data(), &f))) return rc;
        if (dev && nd < dev_cap) dev[nd] = f;
        ++nd;
    }
    if (prm->max_iter > 0) {
        std::memcpy(W, Wu.data(), sizeof(double) * Wu.size());
        std::memcpy(H, Hu.data(), sizeof(double) * Hu.size());
    } else if (!left) {
        std::memcpy(W, Wi.data(), sizeof(double) * Wi.size());
    } else {
        for (int k = 0; k < r; ++k)
            for (int64_t j = 0; j < p; ++j) H[(size_t)k * p + j] = Wi[(size_t)j * r + k];
    }
    if (n_iter) *n_iter = prm->max_iter;
    if (n_dev) *n_dev = nd;
    return NMFG_OK;
}

// One solve from host buffers: V (and lsNMF's weight) up, the iterations,
// W, H (and the offset) back.
int run_host(nmfg_ctx *ctx, const nmfg_params *prm, const double *V, int64_t n, int64_t p,
             int64_t ldv, const double *weight, int64_t ldw, double *W, double *H, int32_t r,
             double *offset, int32_t offset_in, int64_t *n_iter, double *dev, int64_t dev_cap,
             int64_t *n_dev) {
    if (!ctx) return fail(NMFG_EINVAL, "NULL context");
    if (offset_in && !offset) return fail(NMFG_EINVAL, "offset_in set with a NULL offset");
    int rc = check_host_shapes(V, n, p, W, H, r);
    if (rc) return rc;
    if (!prm) return fail(NMFG_EINVAL, "params is NULL");
    if (ldv < p) return fail(NMFG_EINVAL, "ldv < p");
    if (prm->algo == NMFG_SNMF_R || prm->algo == NMFG_SNMF_L)
        return run_snmf(ctx, prm, V, n, p, ldv, W, H, r, n_iter, dev, dev_cap, n_dev);
    if ((rc = validate_params(prm))) return rc;
    if (ldv < p) return fail(NMFG_EINVAL, "ldv < p");
    if (prm->algo == NMFG_LSNMF && !weight)
        return fail(NMFG_EINVAL, "method='lsnmf' requires a `weight` matrix of the same shape as V");
    Guard g(ctx);
    Scoped sc;
    if ((rc = nmfg_matrix_create(ctx, n, p, prm->precision, &sc.m))) return rc;
    if ((rc = nmfg_matrix_upload(sc.m, V, ldv))) return rc;
    if ((rc = nmfg_solve_create(sc.m, r, prm, &sc.s))) return rc;
    if ((rc = nmfg_solve_set_factors(sc.s, W, H))) return rc;
    nmfg_solve *s = sc.s;
    if (s->kind == K_LSNMF && (rc = nmfg_solve_set_weight(s, weight, ldw))) return rc;
    const bool with_off = s->kind == K_OFFSET && offset;
    if (with_off && offset_in && (rc = nmfg_solve_set_offset(s, offset))) return rc;
    if (prm->stop != NMFG_STOP_MAX_ITER) {
        if ((rc = run_with_objective(s, ctx->stream))) return rc;
    } else {
        s->n_iter = prm->max_iter;
        if ((rc = enqueue_fixed(s, ctx->stream))) return rc;
    }
    if ((rc = nmfg_solve_fetch(s, W, H, n_iter, dev, dev_cap, n_dev))) return rc;
    return with_off ? nmfg_solve_get_offset(s, offset) : NMFG_OK;
}

}  // namespace

extern "C" {

int nmfg_fcnnls(const double *ctc, const double *cta, int32_t l, int64_t p, double eps, double *K) {
    if (!ctc || !cta || !K) return fail(NMFG_EINVAL, "NULL argument");
    if (l < 1 || l > 32 || p < 0) return fail(NMFG_EINVAL, "fcnnls: l must be in 1..32, p >= 0");
    return fcnnls_normal(ctc, cta, l, p, eps, K) ? fail(NMFG_EINVAL, "fcnnls failed") : NMFG_OK;
}

int nmfg_run(nmfg_ctx *ctx, const nmfg_params *prm, const double *V, int64_t n, int64_t p,
             int64_t ldv, double *W, double *H, int32_t r, int64_t *n_iter, double *dev,
             int64_t dev_cap, int64_t *n_dev) {
    return run_host(ctx, prm, V, n, p, ldv, nullptr, 0, W, H, r, nullptr, 0, n_iter, dev, dev_cap,
                    n_dev);
}

int nmfg_run_offset(nmfg_ctx *ctx, const nmfg_params *prm, const double *V, int64_t n, int64_t p,
                    int64_t ldv, double *W, double *H, int32_t r, double *offset,
                    int32_t offset_in, int64_t *n_iter, double *dev, int64_t dev_cap,
                    int64_t *n_dev) {
    return run_host(ctx, prm, V, n, p, ldv, nullptr, 0, W, H, r, offset, offset_in, n_iter, dev,
                    dev_cap, n_dev);
}

int nmfg_run_device_noalloc(nmfg_ctx *ctx, const nmfg_params *prm, const double *V, int64_t n,
                            int64_t p, int64_t ldv, double *W, double *H, int64_t ldh, int32_t r,
                            int64_t *n_iter, double *dev, int64_t dev_cap, int64_t *n_dev) {
    if (!ctx || !prm || !V || !W || !H) return fail(NMFG_EINVAL, "NULL argument");
    if (n < 1 || p < 1 || r < 1) return fail(NMFG_EINVAL, "need n, p, r >= 1");
    if (n > (int64_t)1 << 30 || p > (int64_t)1 << 30) return fail(NMFG_EINVAL, "V too large");
    if (ldv < p || ldv % LDV_ALIGN != 0)
        return fail(NMFG_EINVAL, "ldv must be >= p and a multiple of " + std::to_string(LDV_ALIGN));
    if (ldh < p) return fail(NMFG_EINVAL, "ldh < p");
    int rc = validate_params(prm);
    if (rc) return rc;
    const int kind = kind_of(prm->algo);
    if (kind != K_LEE && kind != K_BRUNET && kind != K_HALS && kind != K_EHALS && kind != K_DNMF &&
        kind != K_NSNMF)
        return fail(NMFG_EUNSUPPORTED,
                    "the no-allocation entry runs lee, brunet, hals, ehals, dnmf and nsNMF");
    if (prm->precision != NMFG_FP64)
        return fail(NMFG_EUNSUPPORTED, "the no-allocation entry takes V as FP64 (NMFG_FP64)");
    if (ctx->device >= kMaxDevices) return fail(NMFG_EUNSUPPORTED, "device index too large");
    NoallocDevice &nd = g_noalloc[ctx->device];
    if (nd.err != hipSuccess || !nd.ptr)
        return fail(NMFG_EDEVICE, std::string("no-allocation scratch unavailable: ") +
                                      hipGetErrorString(nd.err));
    Guard g(ctx);
    std::lock_guard<std::mutex> lk(nd.mu);
    hipStream_t st = ctx->stream;
    // V as given (the layout the kernels read: row stride a multiple of 128,
    // padding columns finite), the solve's buffers in the scratch array
    nmfg_matrix m{};
    m.ctx = ctx;
    m.n = n;
    m.p = p;
    m.ldv = ldv;
    m.form = NMFG_FP64;
    m.precision = VP_F64;
    m.V = const_cast<double *>(V);
    nmfg_solve s{};
    if ((rc = solve_setup(&m, r, prm, kind, 1, &s))) return rc;
    const size_t need = solve_bind(&s, nullptr);
    if (need > NOALLOC_SCRATCH_BYTES)
        return fail(NMFG_ENOMEM, "the solve needs " + std::to_string(need) +
                                     " bytes of device scratch; the no-allocation entry has " +
                                     std::to_string(NOALLOC_SCRATCH_BYTES));
    solve_bind(&s, static_cast<char *>(nd.ptr));
    s.no_graph = true;
    HIPCHK(solve_init_buffers(&s));
    FactorPack fp{W, H, n, p, ldh, s.r, s.W0, s.H0, s.R, ldv, 1};
    HIPCHK(launch_factor_pack(fp, st));
    s.have_factors = true;
    if (prm->stop != NMFG_STOP_MAX_ITER) {
        if ((rc = run_with_objective(&s, st))) return rc;
    } else {
        s.n_iter = prm->max_iter;
        if ((rc = enqueue_fixed(&s, st))) return rc;
    }
    fp.sW = s.W;
    fp.sH = s.H;
    fp.to_solve = 0;
    HIPCHK(launch_factor_pack(fp, st));
    HIPCHK(hipStreamSynchronize(st));
    if (n_iter) *n_iter = s.n_iter;
    if (n_dev) *n_dev = (int64_t)s.devs.size();
    if (dev)
        for (int64_t i = 0; i < (int64_t)s.devs.size() && i < dev_cap; ++i) dev[i] = s.devs[i];
    return NMFG_OK;
}

int nmfg_run_lsnmf(nmfg_ctx *ctx, const nmfg_params *prm, const double *V, int64_t n, int64_t p,
                   int64_t ldv, const double *weight, int64_t ldw, double *W, double *H, int32_t r,
                   int64_t *n_iter, double *dev, int64_t dev_cap, int64_t *n_dev) {
    if (!weight) return fail(NMFG_EINVAL, "NULL weight");
    return run_host(ctx, prm, V, n, p, ldv, weight, ldw, W, H, r, nullptr, 0, n_iter, dev, dev_cap,
                    n_dev);
}

}  // extern "C"

// ===========================================================================
// Rank sweeps: `count` solves of uneven ranks on one V (the restarts x ranks
// of a consensus rank sweep, examples/build_scaling_notebook.py:157-208),
// advanced by one replay of one captured HIP graph.  The members are grouped
// by rank into ragged batched solves (each group padded to one template
// rank, each member keeping its own rank); every kernel launch of a group
// covers all its members, and the groups' launches sit on parallel branches
// of the graph.  The plan (grouping, branches, split counts) is chosen by
// timing candidates on the device (autotuner) and cached per shape.
// ===========================================================================
struct nmfg_sweep {
    nmfg_matrix *m = nullptr;
    nmfg_params prm{};
    std::vector<int> ranks;
    struct Group {
        nmfg_solve *s = nullptr;
        std::vector<int> members;  // sweep member indices, in batch order
    };
    std::vector<Group> groups;
    int nstreams = 1;
    std::vector<hipStream_t> streams;  // branch streams (the first one starts the capture)
    hipGraphExec_t graph = nullptr;
    bool have_factors = false;
    std::string plan;     // description of the chosen plan (JSON)
    double tuned_ms = 0;  // autotuner: time of the chosen plan's tuning run
};

namespace {

// Grouping of sorted member ranks: the upper rank of each group.
// kind 0: one group per template rank; 1: bounds 4, 8, 12, 16, 24, 32;
// 2: bounds 8, 16, 32; 3: a single group.
std::vector<int> group_bounds(int plan, int rmax) {
    std::vector<int> b;
    if (plan == 0) {
        for (int r = 1; r <= rmax; ++r) {
            const RankKernels *K = kernels_for_rank_f64(r);
            if (K && (b.empty() || b.back() != K->R)) b.push_back(K->R);
        }
    } else if (plan == 1) {
        b = {4, 8, 12, 16, 24, 32};
    } else if (plan == 2) {
        b = {8, 16, 32};
    } else {
        b = {32};
    }
    return b;
}

struct SweepPlan {
    int grouping = 0;  // group_bounds plan
    int nstreams = 1;
    int split_scale = 0;  // 0: occupancy rules; 1: KL column splits doubled; -1: halved
    double tuned_ms = 0;  // the plan's time in the tuning run (0: not tuned)
};

std::mutex g_tune_mu;
std::map<std::string, SweepPlan> g_tune_cache;

std::string sweep_key(const nmfg_matrix *m, const nmfg_params *prm, const std::vector<int> &ranks) {
    std::vector<int> r = ranks;
    std::sort(r.begin(), r.end());
    std::string k = std::to_string(m->n) + "x" + std::to_string(m->p) + ":f" +
                    std::to_string(m->form) + ":a" + std::to_string(prm->algo) + ":";
    for (int x : r) k += std::to_string(x) + ",";
    return k;
}

void sweep_clear(nmfg_sweep *sw) {
    if (sw->graph) (void)hipGraphExecDestroy(sw->graph);
    sw->graph = nullptr;
    for (auto &g : sw->groups)
        if (g.s) {
            solve_free(g.s);
            delete g.s;
        }
    sw->groups.clear();
    for (size_t i = 1; i < sw->streams.size(); ++i) (void)hipStreamDestroy(sw->streams[i]);
    sw->streams.clear();
}

// Builds the groups of a plan (with iteration count max_iter) and captures
// the graph of one full solve of every member.
int sweep_build(nmfg_sweep *sw, const SweepPlan &pl, int64_t max_iter) {
    sweep_clear(sw);
    nmfg_matrix *m = sw->m;
    hipStream_t st0 = m->ctx->stream;
    const int rmax = *std::max_element(sw->ranks.begin(), sw->ranks.end());
    const std::vector<int> bounds = group_bounds(pl.grouping, rmax);
    std::vector<std::vector<int>> members(bounds.size());
    for (int i = 0; i < (int)sw->ranks.size(); ++i)
        for (size_t g = 0; g < bounds.size(); ++g)
            if (sw->ranks[i] <= bounds[g]) {
                members[g].push_back(i);
                break;
            }
    nmfg_params prm = sw->prm;
    prm.max_iter = max_iter;
    const int kind = kind_of(prm.algo);
    int rc = NMFG_OK;
    for (size_t g = 0; g < bounds.size() && !rc; ++g) {
        if (members[g].empty()) continue;
        std::vector<int> mr;
        int gmax = 0;
        for (int i : members[g]) {
            mr.push_back(sw->ranks[i]);
            gmax = std::max(gmax, sw->ranks[i]);
        }
        nmfg_sweep::Group grp;
        grp.members = members[g];
        // the split counts the occupancy rules pick for this group, scaled
        nmfg_solve probe;
        SolveOverrides ov;
        if (pl.split_scale != 0 && !(rc = solve_setup(m, gmax, &prm, kind, (int)mr.size(), &probe))) {
            if (probe.kl_split > 1 || pl.split_scale > 0)
                ov.kl_split = std::max(1, pl.split_scale > 0 ? 2 * probe.kl_split : probe.kl_split / 2);
        }
        if (!rc)
            rc = solve_create_kind(m, gmax, &prm, kind, (int)mr.size(), &grp.s, mr.data(),
                                   pl.split_scale != 0 ? &ov : nullptr);
        if (!rc) sw->groups.push_back(grp);
    }
    if (rc) return rc;
    const int ns = std::max(1, std::min<int>(pl.nstreams, (int)sw->groups.size()));
    sw->nstreams = ns;
    sw->streams.assign(1, st0);
    for (int i = 1; i < ns; ++i) {
        hipStream_t x = nullptr;
        HIPCHK(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
        sw->streams.push_back(x);
    }
    // one graph: fork to the branch streams, each group's whole solve on
    // its branch, join
    hipGraph_t graph = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins(ns, nullptr);
    HIPCHK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    for (int i = 1; i < ns; ++i) HIPCHK(hipEventCreateWithFlags(&joins[i], hipEventDisableTiming));
    hipError_t e = hipStreamBeginCapture(st0, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess && ns > 1) {
        e = hipEventRecord(fork, st0);
        for (int i = 1; i < ns && e == hipSuccess; ++i) e = hipStreamWaitEvent(sw->streams[i], fork, 0);
    }
    for (size_t g = 0; g < sw->groups.size() && e == hipSuccess && !rc; ++g)
        rc = enqueue_fixed(sw->groups[g].s, sw->streams[g % ns]);
    for (int i = 1; i < ns && e == hipSuccess && !rc; ++i) {
        e = hipEventRecord(joins[i], sw->streams[i]);
        if (e == hipSuccess) e = hipStreamWaitEvent(st0, joins[i], 0);
    }
    const hipError_t ee = hipStreamEndCapture(st0, &graph);
    if (e == hipSuccess) e = ee;
    if (e == hipSuccess && !rc) e = hipGraphInstantiate(&sw->graph, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipEventDestroy(fork);
    for (int i = 1; i < ns; ++i) (void)hipEventDestroy(joins[i]);
    if (rc) return rc;
    if (e != hipSuccess) {
        sw->graph = nullptr;
        (void)hipGetLastError();
        return fail(NMFG_EDEVICE, std::string("sweep graph: ") + hipGetErrorString(e));
    }
    return NMFG_OK;
}

// Milliseconds of one replay of the built graph (after a warm-up replay),
// from uniform factors in [0.1, 1).
int sweep_time(nmfg_sweep *sw, double *ms) {
    hipStream_t st = sw->m->ctx->stream;
    for (auto &g : sw->groups) {
        nmfg_solve *s = g.s;
        size_t wl = 0, hl = 0;
        for (int i : g.members) {
            wl += (size_t)sw->m->n * sw->ranks[i];
            hl += (size_t)sw->ranks[i] * sw->m->p;
        }
        std::vector<double> w(wl), h(hl);
        uint64_t x = 88172645463325252ull;
        auto next = [&x] {
            x ^= x << 13;
            x ^= x >> 7;
            x ^= x << 17;
            return 0.1 + 0.9 * (double)(x >> 11) / 9007199254740992.0;
        };
        for (double &v : w) v = next();
        for (double &v : h) v = next();
        int rc = pack_factors(s, w.data(), h.data(), s->W0, s->H0);
        if (rc) return rc;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    hipError_t e = hipGraphLaunch(sw->graph, st);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    if (e == hipSuccess) e = hipGraphLaunch(sw->graph, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return fail(NMFG_EDEVICE, std::string("sweep timing: ") + hipGetErrorString(e));
    *ms = t;
    return NMFG_OK;
}

std::string plan_json(const nmfg_sweep *sw, const SweepPlan &pl, bool tuned) {
    std::string j = "{\"grouping\": " + std::to_string(pl.grouping) + ", \"branches\": " +
                    std::to_string(sw->nstreams) + ", \"split_scale\": " +
                    std::to_string(pl.split_scale) + ", \"tuned\": " + (tuned ? "true" : "false") +
                    ", \"groups\": [";
    for (size_t g = 0; g < sw->groups.size(); ++g) {
        const nmfg_solve *s = sw->groups[g].s;
        j += std::string(g ? ", " : "") + "{\"R\": " + std::to_string(s->R) + ", \"members\": " +
             std::to_string(s->batch) + ", \"kl_split\": " + std::to_string(s->kl_split) +
             ", \"row_split\": " + std::to_string(s->nsplit) + "}";
    }
    return j + "]}";
}

}  // namespace

extern "C" {

int nmfg_sweep_create(nmfg_matrix *m, const int32_t *ranks, int32_t count, const nmfg_params *prm,
                      int32_t tune, nmfg_sweep **out) {
    if (!m || !ranks || !prm || !out) return fail(NMFG_EINVAL, "NULL argument");
    if (count < 1 || count > 65535) return fail(NMFG_EINVAL, "count must be in 1..65535");
    int rc = validate_params(prm);
    if (rc) return rc;
    if (prm->stop != NMFG_STOP_MAX_ITER)
        return fail(NMFG_EUNSUPPORTED, "a sweep runs a fixed number of iterations");
    const int kind = kind_of(prm->algo);
    if (kind != K_LEE && kind != K_BRUNET && kind != K_HALS && kind != K_EHALS && kind != K_DNMF)
        return fail(NMFG_EUNSUPPORTED, "sweeps run lee, brunet, hals, ehals and dnmf");
    for (int i = 0; i < count; ++i)
        if (ranks[i] < 1 || ranks[i] > max_supported_rank())
            return fail(NMFG_EINVAL, "sweep rank out of range 1.." + std::to_string(max_supported_rank()));
    Guard g(m->ctx);
    nmfg_sweep *sw = new nmfg_sweep();
    sw->m = m;
    sw->prm = *prm;
    sw->ranks.assign(ranks, ranks + count);
    const std::string key = sweep_key(m, prm, sw->ranks);
    SweepPlan best;
    bool have = false, tuned = false;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            best = it->second;
            have = tuned = true;
            sw->tuned_ms = best.tuned_ms;
        }
    }
    if (!have && tune) {
        // time each candidate plan on a short run (10 iterations) and keep
        // the fastest
        const int64_t titer = std::min<int64_t>(10, std::max<int64_t>(1, prm->max_iter));
        double best_ms = 1e300;
        const SweepPlan cands[] = {{0, 1, 0}, {0, 4, 0}, {1, 4, 0}, {1, 1, 0}, {2, 2, 0},
                                   {3, 1, 0}, {1, 4, 1}, {1, 4, -1}};
        for (const SweepPlan &c : cands) {
            if (c.split_scale != 0 && kind != K_BRUNET) continue;
            double ms = 0;
            if ((rc = sweep_build(sw, c, titer)) || (rc = sweep_time(sw, &ms))) break;
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }
        sweep_clear(sw);
        if (rc) {
            delete sw;
            return rc;
        }
        best.tuned_ms = best_ms;
        sw->tuned_ms = best_ms;
        tuned = true;
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = best;
    } else if (!have) {
        best = SweepPlan{1, 4, 0};
    }
    if ((rc = sweep_build(sw, best, prm->max_iter))) {
        sweep_clear(sw);
        delete sw;
        return rc;
    }
    sw->plan = plan_json(sw, best, tuned);
    *out = sw;
    return NMFG_OK;
}

int nmfg_sweep_set_factors(nmfg_sweep *sw, const double *const *W0, const double *const *H0) {
    if (!sw || !W0 || !H0) return fail(NMFG_EINVAL, "NULL argument");
    Guard g(sw->m->ctx);
    const int64_t n = sw->m->n, p = sw->m->p;
    for (auto &grp : sw->groups) {
        size_t wl = 0, hl = 0;
        for (int i : grp.members) {
            if (!W0[i] || !H0[i]) return fail(NMFG_EINVAL, "NULL factor of a sweep member");
            wl += (size_t)n * sw->ranks[i];
            hl += (size_t)sw->ranks[i] * p;
        }
        std::vector<double> w(wl), h(hl);
        size_t ow = 0, oh = 0;
        for (int i : grp.members) {
            const size_t a = (size_t)n * sw->ranks[i], b = (size_t)sw->ranks[i] * p;
            std::memcpy(&w[ow], W0[i], sizeof(double) * a);
            std::memcpy(&h[oh], H0[i], sizeof(double) * b);
            ow += a;
            oh += b;
        }
        int rc = pack_factors(grp.s, w.data(), h.data(), grp.s->W0, grp.s->H0);
        if (rc) return rc;
    }
    sw->have_factors = true;
    return NMFG_OK;
}

int nmfg_sweep_launch(nmfg_sweep *sw) {
    if (!sw) return fail(NMFG_EINVAL, "NULL argument");
    if (!sw->have_factors) return fail(NMFG_EINVAL, "set the members' W0/H0 first");
    Guard g(sw->m->ctx);
    HIPCHK(hipGraphLaunch(sw->graph, sw->m->ctx->stream));
    return NMFG_OK;
}

int nmfg_sweep_fetch(nmfg_sweep *sw, double *const *W, double *const *H) {
    if (!sw || !W || !H) return fail(NMFG_EINVAL, "NULL argument");
    Guard g(sw->m->ctx);
    HIPCHK(hipStreamSynchronize(sw->m->ctx->stream));
    const int64_t n = sw->m->n, p = sw->m->p;
    for (auto &grp : sw->groups) {
        size_t wl = 0, hl = 0;
        for (int i : grp.members) {
            wl += (size_t)n * sw->ranks[i];
            hl += (size_t)sw->ranks[i] * p;
        }
        std::vector<double> w(wl), h(hl);
        int rc = unpack_factors(grp.s, grp.s->W, grp.s->H, w.data(), h.data());
        if (rc) return rc;
        size_t ow = 0, oh = 0;
        for (int i : grp.members) {
            const size_t a = (size_t)n * sw->ranks[i], b = (size_t)sw->ranks[i] * p;
            if (W[i]) std::memcpy(W[i], &w[ow], sizeof(double) * a);
            if (H[i]) std::memcpy(H[i], &h[oh], sizeof(double) * b);
            ow += a;
            oh += b;
        }
    }
    return NMFG_OK;
}

int nmfg_sweep_plan(const nmfg_sweep *sw, char *buf, int64_t cap, double *tuned_ms) {
    if (!sw) return fail(NMFG_EINVAL, "NULL argument");
    if (tuned_ms) *tuned_ms = sw->tuned_ms;
    if (buf && cap > 0) {
        const size_t k = std::min<size_t>(sw->plan.size(), (size_t)cap - 1);
        std::memcpy(buf, sw->plan.data(), k);
        buf[k] = 0;
    }
    return (int64_t)sw->plan.size() < cap ? NMFG_OK : fail(NMFG_EINVAL, "plan buffer too small");
}

int nmfg_sweep_destroy(nmfg_sweep *sw) {
    if (!sw) return NMFG_OK;
    Guard g(sw->m->ctx);
    (void)hipStreamSynchronize(sw->m->ctx->stream);
    sweep_clear(sw);
    delete sw;
    return NMFG_OK;
}

// ---------------------------------------------------------------------------
// Step groups: concurrent solves advanced together.  A step enqueues one
// chunk (STOP_CHECK_CHUNK iterations) of every member, the members spread
// over the group's branch streams (forked from and joined to the context
// stream, so the device runs them side by side), then gathers every
// member's stop flag with one kernel and reads them back with one copy and
// one synchronisation.  Members that stopped (or reached max_iter) get their
// finishing work and leave; new members join at the next step.  A member's
// kernels and their order are those of its own solve run alone, so its
// results are the same bits.
// ---------------------------------------------------------------------------
constexpr int GROUP_MAX_BRANCHES = 8;

struct nmfg_group {
    nmfg_ctx *ctx = nullptr;
    int nbranch = 1;
    hipStream_t br[GROUP_MAX_BRANCHES] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[GROUP_MAX_BRANCHES] = {};
    struct Member {
        nmfg_solve *s = nullptr;
        int64_t iter = 0;
        bool begun = false;
        int branch = 0;              // the branch stream all its work runs on
        hipEvent_t ready = nullptr;  // the work queued on the solve's stream before it joined
        ChunkGraph chunk;            // captured on and launched on the group's streams only
    };
    std::mutex mu;                   // pending and changes of members (adds come from other threads)
    std::vector<Member> pending, members;
    int64_t *d_out = nullptr;        // [stop flag, stop iteration] per member and branch (device)
    int64_t *h_out = nullptr;        // the same, pinned host memory
    bool inflight[GROUP_MAX_BRANCHES] = {};           // a round queued and not yet taken
    std::vector<nmfg_solve *> round[GROUP_MAX_BRANCHES];  // the members of that round, in order
    int64_t steps = 0, member_steps = 0;
    int max_members = 0;
    int chunks = 4;                  // most chunks a member runs per step (NMFG_GROUP_CHUNKS)
};

namespace {

void group_release(nmfg_group *g) {
    for (int b = 0; b < GROUP_MAX_BRANCHES; ++b) {
        if (g->br[b]) (void)hipStreamDestroy(g->br[b]);
        if (g->join[b]) (void)hipEventDestroy(g->join[b]);
    }
    if (g->fork) (void)hipEventDestroy(g->fork);
    for (auto *v : {&g->pending, &g->members})
        for (auto &mb : *v) {
            if (mb.ready) (void)hipEventDestroy(mb.ready);
            if (mb.chunk.exec) (void)hipGraphExecDestroy(mb.chunk.exec);
            if (mb.s) objrun_disarm(mb.s);
        }
    if (g->d_out) ctx_release(g->ctx, g->d_out);
    if (g->h_out) (void)hipHostFree(g->h_out);
}

// A fixed-iteration member: the next chunk of plain iterations.
// Relative cost of one chunk of a group member: V's bytes times the
// template rank, the KL passes about half again as costly per element.
double group_chunk_cost(const nmfg_solve *s) {
    const double eb = s->m->form == NMFG_FP64 ? 8.0 : 4.0;
    return (double)s->m->n * (double)s->m->ldv * eb * (double)s->R * (kl_kind(s->kind) ? 1.5 : 1.0);
}

int fixed_chunk(nmfg_solve *s, hipStream_t st, int64_t *iter) {
    const int64_t chunk_end = std::min<int64_t>(s->prm.max_iter, *iter + STOP_CHECK_CHUNK);
    int rc;
    while (*iter < chunk_end) {
        *iter += 1;
        if ((rc = enqueue_iteration(s, st, *iter))) return rc;
    }
    return NMFG_OK;
}

}  // namespace

int nmfg_group_create(nmfg_ctx *ctx, int32_t branches, nmfg_group **out) {
    if (!ctx || !out) return fail(NMFG_EINVAL, "NULL argument");
    if (branches < 1 || branches > GROUP_MAX_BRANCHES)
        return fail(NMFG_EINVAL, "branches must be 1.." + std::to_string(GROUP_MAX_BRANCHES));
    Guard gd(ctx);
    nmfg_group *g = new nmfg_group();
    g->ctx = ctx;
    g->nbranch = branches;
    if (const char *e = std::getenv("NMFG_GROUP_CHUNKS")) g->chunks = std::max(1, std::atoi(e));
    hipError_t e = hipEventCreateWithFlags(&g->fork, hipEventDisableTiming);
    for (int b = 0; b < branches && e == hipSuccess; ++b) {
        e = hipStreamCreateWithFlags(&g->br[b], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&g->join[b], hipEventDisableTiming);
    }
    if (e == hipSuccess)
        e = hipHostMalloc((void **)&g->h_out, sizeof(int64_t) * 2 * GROUP_MAX * GROUP_MAX_BRANCHES);
    if (e != hipSuccess) {
        group_release(g);
        delete g;
        return fail(NMFG_EDEVICE, std::string("group: ") + hipGetErrorString(e));
    }
    int rc = ctx_alloc(ctx, sizeof(int64_t) * 2 * GROUP_MAX * GROUP_MAX_BRANCHES, (void **)&g->d_out);
    if (rc) {
        group_release(g);
        delete g;
        return rc;
    }
    *out = g;
    return NMFG_OK;
}

int nmfg_group_add(nmfg_group *g, nmfg_solve *s) {
    if (!g || !s) return fail(NMFG_EINVAL, "NULL argument");
    if (!s->have_factors) return fail(NMFG_EINVAL, "set W0/H0 first");
    if (s->kind == K_LSNMF && !s->wgt)
        return fail(NMFG_EINVAL, "method='lsnmf' requires a `weight` matrix of the same shape as V");
    if (s->m->ctx->device != g->ctx->device)
        return fail(NMFG_EINVAL, "the solve is on another device than the group");
    nmfg_group::Member mb;
    mb.s = s;
    {
        Guard gd(s->m->ctx);
        hipError_t e = hipEventCreateWithFlags(&mb.ready, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(mb.ready, s->m->ctx->stream);
        if (e != hipSuccess) {
            if (mb.ready) (void)hipEventDestroy(mb.ready);
            return fail(NMFG_EDEVICE, std::string("group add: ") + hipGetErrorString(e));
        }
    }
    std::lock_guard<std::mutex> lk(g->mu);
    for (auto *v : {&g->pending, &g->members})
        for (auto &o : *v)
            if (o.s == s) {
                (void)hipEventDestroy(mb.ready);
                return fail(NMFG_EINVAL, "the solve is already in the group");
            }
    if (g->pending.size() + g->members.size() >= (size_t)GROUP_MAX) {
        (void)hipEventDestroy(mb.ready);
        return fail(NMFG_EINVAL, "the group holds " + std::to_string(GROUP_MAX) + " solves at most");
    }
    g->pending.push_back(mb);
    return NMFG_OK;
}

int nmfg_group_size(nmfg_group *g, int32_t *members) {
    if (!g || !members) return fail(NMFG_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(g->mu);
    *members = (int32_t)(g->pending.size() + g->members.size());
    return NMFG_OK;
}

int nmfg_group_step(nmfg_group *g, nmfg_solve **finished, int32_t cap, int32_t *n_finished) {
    if (!g || !n_finished) return fail(NMFG_EINVAL, "NULL argument");
    *n_finished = 0;
    Guard gd(g->ctx);
    {
        std::lock_guard<std::mutex> lk(g->mu);
        for (auto &mb : g->pending) {
            // a new member joins the branch with the least chunk cost
            double load[GROUP_MAX_BRANCHES] = {};
            for (const auto &o : g->members) load[o.branch] += group_chunk_cost(o.s);
            mb.branch = (int)(std::min_element(load, load + g->nbranch) - load);
            g->members.push_back(mb);
        }
        g->pending.clear();
    }
    const int cnt = (int)g->members.size();
    if (cnt == 0) return NMFG_OK;
    if (!finished || cap < cnt) return fail(NMFG_EINVAL, "finished[] must hold every member");
    g->max_members = std::max(g->max_members, cnt);
    // A round on every idle branch that has members: chunks of each member
    // (a member whose chunk costs a fraction of the branch's costliest
    // member's runs that many chunks back to back, at most g->chunks), then
    // one kernel gathers the branch's stop flags and one copy brings them
    // back, signalled by the branch's event.
    for (int b = 0; b < g->nbranch; ++b) {
        if (g->inflight[b]) continue;
        std::vector<int> idx;
        for (int i = 0; i < cnt; ++i)
            if (g->members[i].branch == b) idx.push_back(i);
        if (idx.empty()) continue;
        hipStream_t bs = g->br[b];
        double cmax = 0.0;
        for (int i : idx) cmax = std::max(cmax, group_chunk_cost(g->members[i].s));
        StopPtrs ptrs{};
        ptrs.count = (int)idx.size();
        g->round[b].clear();
        int rc = NMFG_OK;
        for (size_t k = 0; k < idx.size() && !rc; ++k) {
            nmfg_group::Member &mb = g->members[idx[k]];
            nmfg_solve *s = mb.s;
            const bool fixed = s->prm.stop == NMFG_STOP_MAX_ITER;
            if (!mb.begun) {
                HIPCHK(hipStreamWaitEvent(bs, mb.ready, 0));
                s->n_iter = 0;
                s->devs.clear();
                if (fixed) {
                    // a stop flag that stays 0 (the gather reads it)
                    HIPCHK(launch_stop_reset(s->d_state, s->d_done, s->d_niter, bs));
                    rc = enqueue_start(s, bs);
                } else {
                    rc = objrun_begin(s, bs);
                }
                mb.begun = true;
            }
            const double c = group_chunk_cost(s);
            const int nch = (int)std::max(1.0, std::min((double)g->chunks, std::floor(c > 0.0 ? cmax / c : 1.0)));
            for (int q = 0; q < nch && !rc && mb.iter < s->prm.max_iter; ++q)
                rc = fixed ? fixed_chunk(s, bs, &mb.iter) : objrun_chunk(s, bs, &mb.iter, &mb.chunk);
            ptrs.done[k] = s->d_done;
            ptrs.niter[k] = s->d_niter;
            g->round[b].push_back(s);
        }
        if (rc) return rc;
        int64_t *dout = g->d_out + (size_t)b * 2 * GROUP_MAX;
        int64_t *hout = g->h_out + (size_t)b * 2 * GROUP_MAX;
        HIPCHK(launch_gather_stop(ptrs, dout, bs));
        HIPCHK(hipMemcpyAsync(hout, dout, sizeof(int64_t) * 2 * ptrs.count, hipMemcpyDeviceToHost, bs));
        HIPCHK(hipEventRecord(g->join[b], bs));
        g->inflight[b] = true;
    }
    // Wait until some round in flight is done (polling every branch's event,
    // so that a short round is taken as soon as it ends); take every round
    // done by then.
    bool any_inflight = false;
    for (int b = 0; b < g->nbranch; ++b) any_inflight = any_inflight || g->inflight[b];
    if (!any_inflight) return NMFG_OK;
    bool ready[GROUP_MAX_BRANCHES] = {};
    for (bool got = false; !got;) {
        for (int b = 0; b < g->nbranch; ++b) {
            if (!g->inflight[b]) continue;
            const hipError_t q = hipEventQuery(g->join[b]);
            if (q == hipSuccess) ready[b] = got = true;
            else if (q != hipErrorNotReady) return fail(NMFG_EDEVICE, std::string("group: ") + hipGetErrorString(q));
        }
        if (!got) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    int nf = 0;
    for (int b = 0; b < g->nbranch; ++b) {
        if (!ready[b]) continue;
        g->inflight[b] = false;
        g->steps += 1;
        g->member_steps += (int64_t)g->round[b].size();
        const int64_t *hout = g->h_out + (size_t)b * 2 * GROUP_MAX;
        bool any = false;
        for (size_t k = 0; k < g->round[b].size(); ++k) {
            nmfg_solve *s = g->round[b][k];
            auto it = std::find_if(g->members.begin(), g->members.end(),
                                   [s](const nmfg_group::Member &o) { return o.s == s; });
            nmfg_group::Member &mb = *it;
            const int done = (int)hout[2 * k];
            if (!done && mb.iter < s->prm.max_iter) continue;
            int rc;
            if (s->prm.stop == NMFG_STOP_MAX_ITER) {
                s->n_iter = s->prm.max_iter;
                rc = enqueue_finish(s, g->br[b], s->prm.max_iter);
            } else {
                rc = objrun_finish(s, g->br[b], mb.iter, done, hout[2 * k + 1]);
            }
            if (rc) return rc;
            (void)hipEventDestroy(mb.ready);
            if (mb.chunk.exec) (void)hipGraphExecDestroy(mb.chunk.exec);  // its launches are done
            {
                // nmfg_group_add and nmfg_group_size read the members under mu
                std::lock_guard<std::mutex> lk(g->mu);
                g->members.erase(it);
            }
            finished[nf++] = s;
            any = true;
        }
        if (any) HIPCHK(hipStreamSynchronize(g->br[b]));
    }
    *n_finished = nf;
    return NMFG_OK;
}

int nmfg_group_stats(nmfg_group *g, int64_t *steps, int64_t *member_steps, int32_t *max_members) {
    if (!g) return fail(NMFG_EINVAL, "NULL argument");
    Guard gd(g->ctx);
    if (steps) *steps = g->steps;
    if (member_steps) *member_steps = g->member_steps;
    if (max_members) *max_members = g->max_members;
    return NMFG_OK;
}

int nmfg_group_destroy(nmfg_group *g) {
    if (!g) return NMFG_OK;
    {
        Guard gd(g->ctx);
        (void)hipStreamSynchronize(g->ctx->stream);
        for (int b = 0; b < g->nbranch; ++b)
            if (g->br[b]) (void)hipStreamSynchronize(g->br[b]);
        group_release(g);
    }
    delete g;
    return NMFG_OK;
}

}  // extern "C"

extern "C" {

int nmfg_footprint(const nmfg_params *prm, int64_t n, int64_t p, int32_t r, int32_t nndsvd_width,
                   int64_t *bytes) {
    if (!prm || !bytes) return fail(NMFG_EINVAL, "NULL argument");
    if (n < 1 || p < 1 || r < 1 || nndsvd_width < 0) return fail(NMFG_EINVAL, "bad shape");
    int rc = validate_params(prm);
    if (rc) return rc;
    const int kind = kind_of(prm->algo);
    nmfg_matrix m{};
    m.n = n;
    m.p = p;
    m.ldv = (p + LDV_ALIGN - 1) / LDV_ALIGN * LDV_ALIGN;
    m.form = prm->precision;
    m.precision = storage_of(prm->precision);
    m.vplane = m.precision == VP_B16X2 ? n * m.ldv : 0;
    m.nc = m.form == NMFG_BF16X2 ? b16_rows(n) : 0;
    // V, and the upload's FP64 staging rows for the 4-byte forms (kept idle
    // in the context's cache afterwards)
    size_t total = block_class(v_block_bytes(m.form, n, m.ldv, nullptr));
    if (m.precision != VP_F64) {
        int64_t rows = std::max<int64_t>(1, ((int64_t)256 << 20) / (8 * p));
        if (m.form == NMFG_BF16X2) rows = std::max<int64_t>(8, rows / 8 * 8);
        rows = std::min<int64_t>(n, rows);
        total += block_class(sizeof(double) * (size_t)rows * p);
    }
    nmfg_solve s;
    if ((rc = solve_setup(&m, r, prm, kind, 1, &s))) return rc;
    total += block_class(solve_bind(&s, nullptr));
    if (nndsvd_width > 0) {
        // product solves of the range finder's chunks (widths 32, then the
        // remainder) and its work space: A Q, A^T Q and the LU work, or the
        // plain products' B and C
        nmfg_params pp;
        nmfg_params_init(&pp);
        pp.max_iter = 0;
        pp.precision = prm->precision;
        for (int c0 = 0; c0 < nndsvd_width; c0 += 32) {
            const int kc = std::min(32, nndsvd_width - c0);
            nmfg_solve ps;
            if ((rc = solve_setup(&m, kc, &pp, K_PROD, 1, &ps))) return rc;
            total += block_class(solve_bind(&ps, nullptr));
        }
        const size_t w = (size_t)nndsvd_width;
        const size_t rf = align256(sizeof(double) * (size_t)n * w) + align256(sizeof(double) * (size_t)p * w) +
                          lu_work_bytes(std::max(n, p), nndsvd_width);
        const size_t mul = sizeof(double) * (size_t)(n + p) * w;
        total += block_class(std::max(rf, mul));
    }
    *bytes = (int64_t)total;
    return NMFG_OK;
}

}  // extern "C"
