// mdn_mech.hip -- host side of the gfx950 mechanical-equilibration library: context, system
// upload (vectorize), neighbor-list rebuild, energy / force entry points and the
// device-driven conjugate-gradient minimizer, behind the C ABI of include/mdn_mech.h.
//
// Citations are relative to medyan-5.4.0/ in the reference tree.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mdn_mech.h"
#include "mdn_kernels.h"
#include "mdn_batched.h"

using namespace mdn;

// Scratch of the force evaluation without allocation (mdn_forces_noalloc): the slot vectors,
// the per-bead bonded sums, the excluded-volume pairs that need the numerical fallback and that
// fallback's partial sums. Every device the library runs on holds its own copy for the life of
// the process. The name is a host-side stand-in, not a device address: host code uses the
// address hipGetSymbolAddress returns for the current device (scratch_dev below), never the name.
constexpr size_t MDN_SCRATCH_BYTES = (size_t)MDN_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) unsigned char scratch[MDN_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[MAX_DEVICES] = {};   // device address of `scratch`, per device

// Looks the array's address up on `device` (current) once; every later call returns it.
hipError_t scratch_lookup(int device) {
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) return hipSuccess;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess) g_scratch_dev[device] = static_cast<unsigned char*>(p);
    return e;
}

// stored slots per interaction (filament stretching / bending forces are computed by the gather)
constexpr int SLOTS_PER_TERM[NFF] = {0, 0, 4, 4, 4, 1};

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t want) {
        if (want <= n && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        size_t m = std::max<size_t>(want, 1);
        hipError_t e = hipMalloc(&p, m * sizeof(T));
        if (e == hipSuccess) n = m;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

struct HipErr {
    hipError_t e;
    const char* what;
};

}  // namespace

struct mdn_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool uploaded = false, nl_built = false;

    // system (host copies used by vectorize)
    int n_beads = 0, n_cyl = 0, n_fil = 0;
    std::vector<int> cyl_b1, cyl_b2, cyl_fil, cyl_pos, cyl_type;
    std::vector<char> cyl_minus;
    std::vector<double> cyl_kexvol, cyl_eql;
    std::vector<mdn_boundary_element> elems;
    double bnd_cutoff = 0, rmax = 0, rmin = 0, grid[3] = {0, 0, 0};

    // device tables
    DBuf<int> str_b, bend_b, lnk_b, mot_b, ev_b, bnd_b, bnd_e, lnk_ord, mot_ord;
    std::vector<int> xl_ord[2];   // linkers, motors: device table position -> crosslink index
    DBuf<double> dof_next, dof_new;   // mdn_dof_serialize: the gathered coordinates, new beads
    DBuf<int> dof_src;
    int dof_next_n = -1;              // beads in dof_next (-1: none pending)
    DBuf<double> str_k, str_eql, bend_k, bend_eqt, lnk_k, lnk_eql, lnk_p1, lnk_p2, lnk_sf;
    DBuf<double> mot_k, mot_eql, mot_p1, mot_p2, mot_sf, ev_ks, bnd_krep, bnd_slen;
    DBuf<int> inc_s_start, inc_s, inc_d_start, inc_d, inc_cnt, inc_cursor, bstart, sorted_id, pos;
    DBuf<unsigned char> pmask;
    DBuf<int> bnbr, bterm;   // filament neighbours and term indices per bead (bonded_at_bead)
    DBuf<double> bsum;       // per-bead bonded sums, 3 * n_beads
    int bonded_gather_mode = -1;   // -1: the defaults below, else both of them
    int bg_eval = 1;                // force evaluations use the bonded sums (bead blocks)
    // parameter gradients / tangents: per cylinder the bending term it is the second cylinder of
    // (-1: none), per bending term its cylinder (host), per-pair excluded-volume energies, a vector
    DBuf<int> d_cyl_bend;
    std::vector<int> bend_cyl;
    DBuf<double> ev_U, grad_out, tan_v;
    DBuf<double> hb_v, hb_slots, hb_piece, hb_out;   // mdn_hvp_batch: per-product tangents, slots, pieces, results
    DBuf<double> piece;   // NBOND * 3 * n_beads values (FP64; the mixed path uses the first half as float)
    // cylinder tables for the NL
    DBuf<int> d_cyl_b1, d_cyl_b2, d_cyl_fil, d_cyl_pos, d_cyl_type;
    DBuf<char> d_cyl_minus;
    DBuf<int> d_fil_ends;      // [2 * n_fil]: first and last cylinder of each filament
    DBuf<double> load_out;     // boundary load forces [2][n_fil][stride]
    DBuf<double> d_cyl_kexvol, d_cyl_eql, d_krep_e, d_slen_e;
    // NL workspace
    DBuf<double> mid;
    DBuf<int> bin, bin_count, bin_start, bin_cursor, members, nl_count, nl_off, pairs, bnd_cnt, bnd_off;
    // the binding-site search's own list (mdn_binding_pairs) and its per-pair counts
    DBuf<double> b_mid, mmid, b_mmid;
    DBuf<int> b_bin, b_bin_count, b_bin_start, b_bin_cursor, b_members, b_count, b_off, b_pairs, b_scount, b_soff;
    DBuf<int> mbin, b_mbin;
    DBuf<float> b_site, b_free;
    DBuf<uint32_t> b_out;
    DBuf<int> nl_err;
    DBuf<unsigned char> cub_tmp;
    long n_pairs = 0;
    int n_bint = 0;
    std::vector<int> bnd_nneighbors;

    // state vectors
    DBuf<double> x[2], f, f2, fp, xh, d[2], xmin[2], scratch_f, ext;   // f, f2: the CG force pair
    DBuf<double> slots;   // 3 values per slot; the mixed path uses the first half as float
    DBuf<double> partial;
    DBuf<int> bad_count, bad_list;
    DBuf<unsigned> ticket;
    DBuf<double> eout;
    DBuf<double> vstat;   // k_vec_stats partials (2 * VSTAT_BLOCKS)
    DBuf<CgState> cg;
    DBuf<double> fallback_partial;
    DBuf<EnDefer> defer;
    DBuf<int> defer_n;
    DBuf<int> fallback_ok;
    CgState* h_cg = nullptr;   // pinned
    double* h_eout = nullptr;  // pinned
    int* h_int = nullptr;      // pinned

    DevSys sys{};
    int cur = 0;   // which x buffer holds the current coordinates

    // batched minimizer (the lead context of a batch owns these)
    DBuf<BatchItem> batch_items;
    DBuf<CgState> batch_states;   // the batch's final states and reports, gathered for one read-back
    DBuf<double> batch_reports;
    DBuf<int> batch_order, batch_queue;
    DBuf<BatchCopy> batch_copy;   // [start copies | start backups | restores], one per system each
    long batch_key = -1;          // (systems, precision, beads) of the tuned schedule
    int batch_nt = 256, batch_sorted = 1, batch_tuned = 0;
    double batch_tune_ms[MDN_BATCH_CANDIDATES] = {};

    bool use_graph = true;   // replay the CG slots as a captured hipGraph (MDN_CG_GRAPH=0: launch them directly)
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};   // per precision (MDN_PREC_*)
    hipGraph_t graph[2] = {nullptr, nullptr};
    long sys_gen = 0, graph_gen[2] = {-1, -1};
    long min_gen[2] = {-1, -1};   // the system generation of the last minimization, per precision
    CgState last_cg{};
};

namespace {

int fail(mdn_ctx* c, int code, const char* fmt, ...) {
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return code;
}

#define HIPCHK(call)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(ctx, MDN_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),  \
                        __FILE__, __LINE__);                                                    \
    } while (0)

#define HIPCHK_CTX(c, call)                                                                     \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail((c), MDN_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),  \
                        __FILE__, __LINE__);                                                    \
    } while (0)

// Device-to-host read ordered after everything queued on the context's stream (a blocking
// hipMemcpy runs on the null stream, which does not wait for the non-blocking context stream).
hipError_t d2h(mdn_ctx* ctx, void* dst, const void* src, size_t bytes) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(ctx->stream);
}

bool call_limit_host(const CgState& S) {
    return (S.energy_call_limit != 0 && S.num_energy_call > S.energy_call_limit) ||
           (S.force_call_limit != 0 && S.num_force_call > S.force_call_limit);
}

template <typename T>
hipError_t upload(mdn_ctx* ctx, DBuf<T>& b, const std::vector<T>& v) {
    hipError_t e = b.ensure(v.size());
    if (e != hipSuccess) return e;
    if (!v.empty()) e = hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream);
    return e;
}

int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }
const int VSTAT_BLOCKS = 1024;   // grid of k_vec_stats

// Block and slot layout of the term / bead grids for the current interaction counts.
void layout(mdn_ctx* ctx) {
    DevSys& s = ctx->sys;
    s.n_beads = ctx->n_beads;
    // the longest-running blocks are dispatched first (excluded volume, crosslinks, boundary,
    // then the bonded terms): on grids larger than the chip holds at once, the cheap bonded
    // blocks then fill in around them instead of delaying them to the kernel's tail
    static const int order[NFF] = {FF_EXVOL, FF_LINKER, FF_MOTOR, FF_BOUNDARY, FF_BEND, FF_STRETCH};
    s.blk[0] = 0;
    for (int k = 0; k < NFF; k++) {
        s.dff[k] = order[k];
        s.fblk[order[k]] = s.blk[k];
        s.blk[k + 1] = s.blk[k] + ceil_div(s.n[order[k]], BLOCK);
    }
    s.slot0[0] = 0;
    for (int ff = 0; ff < NFF; ff++) s.slot0[ff + 1] = s.slot0[ff] + (long)SLOTS_PER_TERM[ff] * s.n[ff];
    s.blk[NFF + 1] = s.blk[NFF] + ceil_div(ctx->n_beads, BLOCK);
    s.nbb = std::max(1, ceil_div(ctx->n_beads, BLOCK));
    s.lsp = s.blk[NFF] + 2 * s.nbb;
    s.blk_nb = s.blk[4];   // excluded volume, linkers, motors, boundary
    // Force evaluations use the bonded sums (bead blocks in the terms kernel's grid recompute the
    // filament terms' forces per bead, which saves writing and re-reading five piece vectors per
    // bead); the minimizer chooses per precision (cg_bonded_sums). The same values either way; a
    // context can fix every choice (bonded_gather_mode).
    s.bonded_gather = 0;   // set per precision on the minimizer's copy (cg_sys)
    ctx->bg_eval = ctx->bonded_gather_mode >= 0 ? ctx->bonded_gather_mode : 1;
}

int bead_blocks(mdn_ctx* ctx) { return std::max(1, ceil_div(ctx->n_beads, BLOCK)); }
int gather_size(mdn_ctx* ctx) { return ctx->n_beads >= GB_LARGE_BEADS ? 128 : 64; }
int gather_blocks(mdn_ctx* ctx) { return std::max(1, ceil_div(ctx->n_beads, gather_size(ctx))); }
// the gather kernels at the context's beads-per-block
template <typename R>
void launch_gather_plain(mdn_ctx* ctx, const DevSys& s, int bmask, const R* slots, long lo, long hi, const double* base,
                         double* out, bool bg) {
    const dim3 g(gather_blocks(ctx)), b(BLOCK);
    const bool l = gather_size(ctx) == 128;
    if (bg && l)
        hipLaunchKernelGGL((k_gather_plain<R, 128, true>), g, b, 0, ctx->stream, s, bmask, slots, lo, hi, base, out);
    else if (bg)
        hipLaunchKernelGGL((k_gather_plain<R, 64, true>), g, b, 0, ctx->stream, s, bmask, slots, lo, hi, base, out);
    else if (l)
        hipLaunchKernelGGL((k_gather_plain<R, 128, false>), g, b, 0, ctx->stream, s, bmask, slots, lo, hi, base, out);
    else
        hipLaunchKernelGGL((k_gather_plain<R, 64, false>), g, b, 0, ctx->stream, s, bmask, slots, lo, hi, base, out);
}
// The terms kernel's grid: every interaction block (pieces: the filament terms write their
// pieces), or the other force fields' blocks followed by the bead blocks of the bonded sums.
int force_blocks(const DevSys& s, bool pieces) { return pieces ? s.blk[NFF] : s.blk_nb + ceil_div(s.n_beads, BLOCK); }
int bead_block0(const DevSys& s, bool pieces) { return pieces ? -1 : s.blk_nb; }
template <typename R>
void launch_cg_gather(mdn_ctx* ctx, const DevSys& s, const CgState* st, const CgBufs& B) {
    const dim3 g(gather_blocks(ctx)), b(BLOCK);
    const bool l = gather_size(ctx) == 128, bg = s.bonded_gather != 0;
    if (bg && l) hipLaunchKernelGGL((k_cg_gather<R, 128, true>), g, b, 0, ctx->stream, s, st, B);
    else if (bg) hipLaunchKernelGGL((k_cg_gather<R, 64, true>), g, b, 0, ctx->stream, s, st, B);
    else if (l) hipLaunchKernelGGL((k_cg_gather<R, 128, false>), g, b, 0, ctx->stream, s, st, B);
    else hipLaunchKernelGGL((k_cg_gather<R, 64, false>), g, b, 0, ctx->stream, s, st, B);
}

void refresh_ptrs(mdn_ctx* ctx) {
    DevSys& s = ctx->sys;
    s.str_b = ctx->str_b.p; s.str_k = ctx->str_k.p; s.str_eql = ctx->str_eql.p;
    s.bend_b = ctx->bend_b.p; s.bend_k = ctx->bend_k.p; s.bend_eqt = ctx->bend_eqt.p;
    s.lnk_b = ctx->lnk_b.p; s.lnk_k = ctx->lnk_k.p; s.lnk_eql = ctx->lnk_eql.p; s.lnk_p1 = ctx->lnk_p1.p;
    s.lnk_p2 = ctx->lnk_p2.p; s.lnk_sf = ctx->lnk_sf.p;
    s.lnk_ord = ctx->lnk_ord.p; s.mot_ord = ctx->mot_ord.p;
    s.mot_b = ctx->mot_b.p; s.mot_k = ctx->mot_k.p; s.mot_eql = ctx->mot_eql.p; s.mot_p1 = ctx->mot_p1.p;
    s.mot_p2 = ctx->mot_p2.p; s.mot_sf = ctx->mot_sf.p;
    s.ev_b = ctx->ev_b.p; s.ev_ks = ctx->ev_ks.p;
    s.bnd_b = ctx->bnd_b.p; s.bnd_e = ctx->bnd_e.p; s.bnd_krep = ctx->bnd_krep.p; s.bnd_slen = ctx->bnd_slen.p;
    s.piece = (void*)ctx->piece.p; s.pmask = ctx->pmask.p;
    s.bnbr = ctx->bnbr.p; s.bterm = ctx->bterm.p; s.bsum = ctx->bsum.p;
    s.pos = ctx->pos.p; s.bstart = ctx->bstart.p; s.sorted_id = ctx->sorted_id.p;
    s.defer = ctx->defer.p; s.defer_n = ctx->defer_n.p;
    s.n_elem = (int)ctx->elems.size();
    for (int e = 0; e < s.n_elem; e++) {
        s.elem[e].kind = ctx->elems[e].kind;
        for (int q = 0; q < 4; q++) s.elem[e].p[q] = ctx->elems[e].p[q];
    }
}

// Size the per-evaluation buffers (slots, partials) for the current layout.
int ensure_eval_buffers(mdn_ctx* ctx) {
    const DevSys& s = ctx->sys;
    // padded by one gather chunk: the gather's chunk loads are unconditional
    HIPCHK(ctx->slots.ensure(3 * (size_t)std::max(1L, s.slot0[NFF]) + GCHUNK_PAD));
    // line-search partials (interaction blocks, then the line search's and the bead kernel's bead
    // blocks; this also covers the evaluation partials), then the CG gather's
    const long nls = (long)s.lsp;
    HIPCHK(ctx->partial.ensure((size_t)(nls + gather_blocks(ctx) + 1) * PSTRIDE));
    HIPCHK(ctx->pos.ensure((size_t)std::max(1L, s.slot0[NFF])));
    HIPCHK(ctx->sorted_id.ensure((size_t)std::max(1L, s.slot0[NFF])));
    HIPCHK(ctx->bad_list.ensure((size_t)std::max(1, s.n[FF_EXVOL])));
    return MDN_OK;
}

// The tile edge of the list kernel: the largest of 8, 4, 2 bins that still gives 1024 tiles (a
// workgroup each), so that small grids keep the chip busy.
void nl_tiles(NLDev& n) {
    for (int T = 8, lt = 3; T >= 2; T /= 2, lt--) {
        n.T = T;
        n.lt = lt;
        long tiles = 1;
        for (int q = 0; q < 3; q++) {
            n.nt[q] = (n.nb[q] + T - 1) / T;
            tiles *= n.nt[q];
        }
        if (tiles >= 1024) break;
    }
}

NLDev nl_view(mdn_ctx* ctx) {
    NLDev n{};
    n.n_cyl = ctx->n_cyl;
    n.cyl_b1 = ctx->d_cyl_b1.p; n.cyl_b2 = ctx->d_cyl_b2.p; n.cyl_fil = ctx->d_cyl_fil.p;
    n.cyl_pos = ctx->d_cyl_pos.p; n.cyl_type = ctx->d_cyl_type.p;
    n.cyl_kexvol = ctx->d_cyl_kexvol.p; n.cyl_eql = ctx->d_cyl_eql.p; n.cyl_minus = ctx->d_cyl_minus.p;
    // setneighborsearchparameters(..., float rMax, float rMin): squares formed in float
    // (HybridNeighborListImpl.h:74-80)
    float rmaxf = (float)ctx->rmax, rminf = (float)ctx->rmin;
    float rmaxsq = rmaxf * rmaxf, rminsq = rminf * rminf;
    n.rmaxsq = rmaxsq;
    n.rminsq = rminsq;
    n.side = 1.125 * std::sqrt((double)rmaxsq);   // HybridNeighborListImpl.cpp:180
    for (int q = 0; q < 3; q++) {
        n.grid[q] = ctx->grid[q];
        n.nb[q] = std::max(1, (int)std::ceil(ctx->grid[q] / n.side));
    }
    n.ta = n.tb = -1;
    nl_tiles(n);
    return n;
}

int exclusive_scan(mdn_ctx* ctx, const int* in, int* out, long n) {
    size_t bytes = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, ctx->stream));
    HIPCHK(ctx->cub_tmp.ensure(bytes));
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp.p, bytes, in, out, (int)n, ctx->stream));
    return MDN_OK;
}

int need_system(mdn_ctx* ctx, bool need_nl) {
    if (!ctx) return MDN_ERR_ARG;
    if (!ctx->uploaded) return fail(ctx, MDN_ERR_STATE, "no system uploaded (mdn_system_upload)");
    if (need_nl && !ctx->nl_built) return fail(ctx, MDN_ERR_STATE, "neighbor lists not built (mdn_nl_build)");
    return MDN_OK;
}

int set_coords_dev(mdn_ctx* ctx, const double* coord) {
    if (coord)
        HIPCHK(hipMemcpyAsync(ctx->x[ctx->cur].p, coord, sizeof(double) * 3 * ctx->n_beads,
                              hipMemcpyHostToDevice, ctx->stream));
    return MDN_OK;
}

// Term forces + numerical fallback for non-finite excluded-volume pairs, into ctx->slots (in
// the arithmetic type R), at device coordinates x.
template <typename R>
int eval_slots_t(mdn_ctx* ctx, const double* x, bool pieces, int bmask, const double* base) {
    const DevSys& s = ctx->sys;
    R* slots = reinterpret_cast<R*>(ctx->slots.p);
    HIPCHK(hipMemsetAsync(ctx->bad_count.p, 0, sizeof(int), ctx->stream));
    if (force_blocks(s, pieces) > 0)
        hipLaunchKernelGGL(k_forces_terms<R>, dim3(force_blocks(s, pieces)), dim3(BLOCK), 0, ctx->stream, s, x, slots,
                           ctx->bad_count.p, ctx->bad_list.p, bead_block0(s, pieces), bmask, base);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(ctx->h_int, ctx->bad_count.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    int nbad = ctx->h_int[0];
    if (nbad == 0) return MDN_OK;
    std::vector<int> bad(nbad);
    HIPCHK(d2h(ctx, bad.data(), ctx->bad_list.p, sizeof(int) * nbad));
    std::sort(bad.begin(), bad.end());
    const int G = 240;
    HIPCHK(ctx->fallback_partial.ensure(6 * G));
    HIPCHK(ctx->fallback_ok.ensure(1));
    for (int t : bad) {
        // forceN (FP64 in both paths), retried once with beads 1 and 2 moved out of plane
        // (CylinderExclVolRepulsion.cpp:1198-1204)
        int ok = 0;
        for (int mv = 0; mv < 2 && !ok; mv++) {
            hipLaunchKernelGGL(k_exvol_forceN_rows, dim3(G), dim3(BLOCK), 0, ctx->stream, s, x, t, mv,
                               ctx->fallback_partial.p);
            hipLaunchKernelGGL(k_exvol_forceN_finish<R>, dim3(1), dim3(64), 0, ctx->stream, s, x, t, mv, G,
                               (const double*)ctx->fallback_partial.p, slots, ctx->fallback_ok.p);
            HIPCHK(hipMemcpyAsync(ctx->h_int, ctx->fallback_ok.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            ok = ctx->h_int[0];
        }
        if (!ok)
            return fail(ctx, MDN_ERR_NONFINITE,
                        "excluded-volume pair %d: force non-finite after the numerical fallback", t);
    }
    return MDN_OK;
}
// pieces: the filament terms' pieces for a gather that reads them; else the bonded sums of
// the pieces selected by bmask on top of base (null: zero) for a gather that starts from them
int eval_slots(mdn_ctx* ctx, const double* x, int prec, bool pieces, int bmask = 3, const double* base = nullptr) {
    return prec == MDN_PREC_MIXED ? eval_slots_t<float>(ctx, x, pieces, bmask, base)
                                  : eval_slots_t<double>(ctx, x, pieces, bmask, base);
}

// force = base + the bonded pieces selected by bmask (bit 0 stretching, bit 1 bending) + the
// slots with id in [lo, hi). bg: the first two terms come as the bonded sums the terms kernel's
// bead blocks wrote (eval_slots without pieces, with the same bmask and base).
int gather(mdn_ctx* ctx, int bmask, long lo, long hi, const double* base, double* out, int prec, bool bg) {
    if (prec == MDN_PREC_MIXED)
        launch_gather_plain<float>(ctx, ctx->sys, bmask, (const float*)ctx->slots.p, lo, hi, base, out, bg);
    else
        launch_gather_plain<double>(ctx, ctx->sys, bmask, (const double*)ctx->slots.p, lo, hi, base, out, bg);
    HIPCHK(hipGetLastError());
    return MDN_OK;
}

void launch_energy(mdn_ctx* ctx, const double* x, int prec) {
    const DevSys& s = ctx->sys;
    int grid = std::max(1, s.blk[NFF]);
    if (prec == MDN_PREC_MIXED)
        hipLaunchKernelGGL(k_energy_report<float>, dim3(grid), dim3(BLOCK), 0, ctx->stream, s, x, ctx->partial.p);
    else
        hipLaunchKernelGGL(k_energy_report<double>, dim3(grid), dim3(BLOCK), 0, ctx->stream, s, x, ctx->partial.p);
    hipLaunchKernelGGL(k_energy_report_fin, dim3(1), dim3(BLOCK), 0, ctx->stream, s, (const double*)ctx->partial.p,
                       ctx->eout.p);
}

// Per-FF energies at device coordinates x into ctx->h_eout[0..7] (synchronous).
int energy_report(mdn_ctx* ctx, const double* x, int prec) {
    launch_energy(ctx, x, prec);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(ctx->h_eout, ctx->eout.p, sizeof(double) * 9, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->h_eout[8] != 0.0)
        return fail(ctx, MDN_ERR_UNSUPPORTED, "%.0f excluded-volume pairs need the Simpson energy (capacity %d)",
                    ctx->h_eout[8], DEFER_CAP);
    return MDN_OK;
}

void fill_report(const double* e, mdn_energy_report* out) {
    for (int ff = 0; ff < NFF; ff++) out->ff[ff] = e[ff];
    out->total = e[6];
}

}  // namespace

// =========================================================================================
// C ABI
// =========================================================================================
extern "C" {

int mdn_abi_version(void) { return MDN_ABI_VERSION; }

const char* mdn_last_error(const mdn_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int mdn_ctx_create(int device, mdn_ctx** out) {
    if (!out) return MDN_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return MDN_ERR_HIP;
    if (device < 0 || device >= n) return MDN_ERR_ARG;
    mdn_ctx* ctx = new mdn_ctx();
    ctx->device = device;
    // MDN_BONDED_GATHER=0 / 1 at context creation fixes whether the filament terms' forces go
    // through per-bead bonded sums (1) or the piece arrays (0), for force evaluations and the
    // minimizer alike; unset: the measured defaults (layout, cg_bonded_sums)
    if (const char* env = getenv("MDN_BONDED_GATHER")) ctx->bonded_gather_mode = atoi(env) != 0 ? 1 : 0;
    // MDN_CG_GRAPH=0 at context creation: the minimizer launches each batch of slots directly
    // instead of replaying a captured graph (the same kernels and decisions; for tools that do
    // not follow graph launches)
    if (const char* env = getenv("MDN_CG_GRAPH")) ctx->use_graph = atoi(env) != 0;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return MDN_ERR_HIP;
    }
    if (scratch_lookup(device) != hipSuccess) {   // the no-allocation entry point's array
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return MDN_ERR_HIP;
    }
    if (hipHostMalloc(&ctx->h_cg, sizeof(CgState)) != hipSuccess ||
        hipHostMalloc(&ctx->h_eout, sizeof(double) * 32) != hipSuccess ||
        hipHostMalloc(&ctx->h_int, sizeof(int) * 16) != hipSuccess) {
        delete ctx;
        return MDN_ERR_HIP;
    }
    if (ctx->ticket.ensure(8) != hipSuccess || ctx->eout.ensure(32) != hipSuccess ||
        ctx->vstat.ensure(2 * VSTAT_BLOCKS) != hipSuccess ||
        ctx->bad_count.ensure(1) != hipSuccess || ctx->cg.ensure(1) != hipSuccess ||
        ctx->nl_err.ensure(1) != hipSuccess || ctx->defer.ensure(DEFER_CAP) != hipSuccess ||
        ctx->defer_n.ensure(1) != hipSuccess) {
        delete ctx;
        return MDN_ERR_HIP;
    }
    (void)hipMemset(ctx->ticket.p, 0, sizeof(unsigned) * 8);
    (void)hipMemset(ctx->defer_n.p, 0, sizeof(int));
    *out = ctx;
    return MDN_OK;
}

int mdn_ctx_destroy(mdn_ctx* ctx) {
    if (!ctx) return MDN_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (int q = 0; q < 2; q++) {
        if (ctx->graph_exec[q]) (void)hipGraphExecDestroy(ctx->graph_exec[q]);
        if (ctx->graph[q]) (void)hipGraphDestroy(ctx->graph[q]);
    }
    DBuf<int>* ib[] = {&ctx->str_b, &ctx->bend_b, &ctx->lnk_b, &ctx->mot_b, &ctx->lnk_ord, &ctx->mot_ord, &ctx->ev_b,
                       &ctx->bnd_b, &ctx->bnd_e,
                       &ctx->inc_s_start, &ctx->inc_s, &ctx->inc_d_start, &ctx->inc_d, &ctx->inc_cnt,
                       &ctx->inc_cursor, &ctx->d_cyl_b1, &ctx->d_cyl_b2, &ctx->d_cyl_fil, &ctx->d_cyl_pos,
                       &ctx->d_cyl_type, &ctx->bin, &ctx->bin_count, &ctx->bin_start, &ctx->bin_cursor,
                       &ctx->members, &ctx->nl_count, &ctx->nl_off, &ctx->pairs, &ctx->bnd_cnt, &ctx->bnd_off,
                       &ctx->nl_err, &ctx->bad_count, &ctx->bad_list, &ctx->fallback_ok, &ctx->bstart,
                       &ctx->sorted_id, &ctx->pos, &ctx->dof_src, &ctx->b_bin, &ctx->b_bin_count,
                       &ctx->b_bin_start, &ctx->b_bin_cursor, &ctx->b_members, &ctx->b_count, &ctx->b_off,
                       &ctx->b_pairs, &ctx->b_scount, &ctx->b_soff};
    ctx->b_mid.release();
    ctx->mmid.release();
    ctx->b_mmid.release();
    ctx->mbin.release();
    ctx->b_mbin.release();
    ctx->b_site.release();
    ctx->b_free.release();
    ctx->b_out.release();
    for (auto* b : ib) b->release();
    DBuf<double>* db[] = {&ctx->str_k, &ctx->str_eql, &ctx->bend_k, &ctx->bend_eqt, &ctx->lnk_k, &ctx->lnk_eql,
                          &ctx->lnk_p1, &ctx->lnk_p2, &ctx->lnk_sf, &ctx->mot_k, &ctx->mot_eql, &ctx->mot_p1,
                          &ctx->mot_p2, &ctx->mot_sf, &ctx->ev_ks, &ctx->bnd_krep, &ctx->bnd_slen,
                          &ctx->d_cyl_kexvol, &ctx->d_cyl_eql, &ctx->d_krep_e, &ctx->d_slen_e, &ctx->mid,
                          &ctx->x[0], &ctx->x[1], &ctx->f, &ctx->f2, &ctx->fp, &ctx->xh, &ctx->d[0], &ctx->d[1],
                          &ctx->xmin[0], &ctx->xmin[1], &ctx->scratch_f, &ctx->ext, &ctx->partial, &ctx->eout,
                          &ctx->vstat, &ctx->fallback_partial, &ctx->dof_next, &ctx->dof_new};
    for (auto* b : db) b->release();
    ctx->d_cyl_minus.release();
    ctx->d_fil_ends.release();
    ctx->load_out.release();
    ctx->pmask.release();
    ctx->bnbr.release();
    ctx->bterm.release();
    ctx->bsum.release();
    ctx->d_cyl_bend.release();
    ctx->ev_U.release();
    ctx->grad_out.release();
    ctx->tan_v.release();
    ctx->hb_v.release();
    ctx->hb_slots.release();
    ctx->hb_piece.release();
    ctx->hb_out.release();
    ctx->piece.release();
    ctx->cub_tmp.release();
    ctx->slots.release();
    ctx->ticket.release();
    ctx->defer.release();
    ctx->defer_n.release();
    ctx->batch_items.release();
    ctx->batch_states.release();
    ctx->batch_reports.release();
    ctx->batch_order.release();
    ctx->batch_queue.release();
    ctx->batch_copy.release();
    ctx->cg.release();
    if (ctx->h_cg) (void)hipHostFree(ctx->h_cg);
    if (ctx->h_eout) (void)hipHostFree(ctx->h_eout);
    if (ctx->h_int) (void)hipHostFree(ctx->h_int);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return MDN_OK;
}

int relayout(mdn_ctx* ctx);

// LinkerStretching / MotorGhostStretching::vectorize (LinkerStretching.cpp:30-57,
// MotorGhostStretching.cpp:30-57): the crosslink tables (beads of both cylinders, k, rest length,
// binding positions) and the static slot incidence, per bead the crosslink slots in ascending
// slot order. The counts are set and the block / slot layout recomputed; the merged bead-sorted
// layout is then stale until relayout (or the next neighbor-list build). Indices are validated
// by the callers.
int set_crosslinks(mdn_ctx* ctx, const mdn_crosslink* linkers, int n_linkers, const mdn_crosslink* motors,
                   int n_motors) {
    DevSys& s = ctx->sys;
    std::vector<int> xb[2];
    std::vector<double> xk[2], xe[2], xp1[2], xp2[2];
    for (int w = 0; w < 2; w++) {
        const int n = w ? n_motors : n_linkers;
        const mdn_crosslink* xl = w ? motors : linkers;
        xb[w].resize(4 * (size_t)n);
        xk[w].resize(n); xe[w].resize(n); xp1[w].resize(n); xp2[w].resize(n);
        for (int j = 0; j < n; j++) {
            const int c1 = xl[j].cyl1, c2 = xl[j].cyl2;
            xb[w][4 * j] = ctx->cyl_b1[c1]; xb[w][4 * j + 1] = ctx->cyl_b2[c1];
            xb[w][4 * j + 2] = ctx->cyl_b1[c2]; xb[w][4 * j + 3] = ctx->cyl_b2[c2];
            xk[w][j] = xl[j].kstr; xe[w][j] = xl[j].eql; xp1[w][j] = xl[j].pos1; xp2[w][j] = xl[j].pos2;
        }
    }
    s.n[FF_LINKER] = n_linkers;
    s.n[FF_MOTOR] = n_motors;
    layout(ctx);
    if (s.slot0[NFF] > 0x7fffffffL) return fail(ctx, MDN_ERR_UNSUPPORTED, "too many interactions");
    // static incidence CSR: per bead, slots of linkers / motors ascending
    std::vector<int> cnt(ctx->n_beads + 1, 0);
    auto each_static = [&](auto&& fn) {
        for (int w = 0; w < 2; w++)
            for (size_t t = 0; t < xk[w].size(); t++)
                for (int q = 0; q < 4; q++) fn(xb[w][4 * t + q], s.slot0[2 + w] + 4L * t + q);
    };
    each_static([&](int b, long) { cnt[b + 1]++; });
    for (int b = 0; b < ctx->n_beads; b++) cnt[b + 1] += cnt[b];
    std::vector<int> inc(std::max(1, cnt[ctx->n_beads]));
    {
        std::vector<int> cur(cnt.begin(), cnt.end() - 1);
        each_static([&](int b, long sl) { inc[cur[b]++] = (int)sl; });   // visited in slot order
    }
    HIPCHK(upload(ctx, ctx->inc_s_start, cnt)); HIPCHK(upload(ctx, ctx->inc_s, inc));
    // The device tables in the threads' order: by the first bead of the first cylinder (stable:
    // ties keep the interaction order); ord[i] is the crosslink's own index (its slots).
    // MDN_XL_ORDER=0 keeps the interaction order (the forces are the same).
    const char* env = getenv("MDN_XL_ORDER");
    const bool sorted = !(env && env[0] == '0');
    std::vector<int> sb[2];
    std::vector<double> sk[2], se[2], sp1[2], sp2[2];
    for (int w = 0; w < 2; w++) {
        std::vector<int>& ord = ctx->xl_ord[w];
        const size_t n = xk[w].size();
        ord.resize(n);
        for (size_t j = 0; j < n; j++) ord[j] = (int)j;
        const std::vector<int>& b = xb[w];
        if (sorted)
            std::stable_sort(ord.begin(), ord.end(), [&b](int i, int j) { return b[4 * (size_t)i] < b[4 * (size_t)j]; });
        sb[w].resize(4 * n); sk[w].resize(n); se[w].resize(n); sp1[w].resize(n); sp2[w].resize(n);
        for (size_t i = 0; i < n; i++) {
            const int t = ord[i];
            for (int q = 0; q < 4; q++) sb[w][4 * i + q] = b[4 * (size_t)t + q];
            sk[w][i] = xk[w][t]; se[w][i] = xe[w][t]; sp1[w][i] = xp1[w][t]; sp2[w][i] = xp2[w][t];
        }
    }
    for (int w = 0; w < 2; w++) {
        std::vector<int> ordu(ctx->xl_ord[w]);
        if (ordu.empty()) ordu.push_back(0);
        HIPCHK(upload(ctx, w ? ctx->mot_ord : ctx->lnk_ord, ordu));
        HIPCHK(upload(ctx, w ? ctx->mot_b : ctx->lnk_b, sb[w]));
        HIPCHK(upload(ctx, w ? ctx->mot_k : ctx->lnk_k, sk[w]));
        HIPCHK(upload(ctx, w ? ctx->mot_eql : ctx->lnk_eql, se[w]));
        HIPCHK(upload(ctx, w ? ctx->mot_p1 : ctx->lnk_p1, sp1[w]));
        HIPCHK(upload(ctx, w ? ctx->mot_p2 : ctx->lnk_p2, sp2[w]));
        HIPCHK((w ? ctx->mot_sf : ctx->lnk_sf).ensure(std::max<size_t>(1, xk[w].size())));
        HIPCHK(hipStreamSynchronize(ctx->stream));   // the host vectors above end with this scope
    }
    return MDN_OK;
}

int mdn_system_upload(mdn_ctx* ctx, const mdn_system_desc* d, const double* coord) {
    if (!ctx || !d) return fail(ctx, MDN_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(ctx->device));
    if (d->n_beads <= 0 || d->n_filaments < 0) return fail(ctx, MDN_ERR_ARG, "bad bead / filament counts");
    if (!coord && ctx->dof_next_n != d->n_beads)
        return fail(ctx, MDN_ERR_ARG, "no coordinates: coord is NULL and mdn_dof_serialize gathered %d beads, not %d",
                    ctx->dof_next_n, d->n_beads);
    if (d->n_boundary > MAXELEM) return fail(ctx, MDN_ERR_UNSUPPORTED, "at most %d boundary elements", MAXELEM);
    ctx->uploaded = ctx->nl_built = false;
    int st = MDN_OK;
    ctx->n_beads = d->n_beads;
    ctx->n_fil = d->n_filaments;
    int n_cyl = 0;
    for (int f = 0; f < d->n_filaments; f++) {
        int nb = d->fil_bead_offset[f + 1] - d->fil_bead_offset[f];
        if (nb < 2) return fail(ctx, MDN_ERR_ARG, "filament %d has %d beads (needs >= 2)", f, nb);
        n_cyl += nb - 1;
    }
    ctx->n_cyl = n_cyl;
    ctx->cyl_b1.assign(n_cyl, 0); ctx->cyl_b2.assign(n_cyl, 0); ctx->cyl_fil.assign(n_cyl, 0);
    ctx->cyl_pos.assign(n_cyl, 0); ctx->cyl_type.assign(n_cyl, 0); ctx->cyl_minus.assign(n_cyl, 0);
    ctx->cyl_kexvol.assign(n_cyl, 0); ctx->cyl_eql.assign(n_cyl, 0);
    // FilamentStretching::vectorize (FilamentStretching.cpp:26-45)
    std::vector<int> str_b(2 * (size_t)n_cyl);
    std::vector<double> str_k(n_cyl), str_eql(n_cyl);
    int c = 0;
    for (int f = 0; f < d->n_filaments; f++) {
        int off = d->fil_bead_offset[f], nb = d->fil_bead_offset[f + 1] - off;
        for (int k = 0; k < nb - 1; k++, c++) {
            int b1 = d->fil_beads[off + k], b2 = d->fil_beads[off + k + 1];
            if (b1 < 0 || b1 >= d->n_beads || b2 < 0 || b2 >= d->n_beads)
                return fail(ctx, MDN_ERR_ARG, "filament %d bead index out of range", f);
            ctx->cyl_b1[c] = b1; ctx->cyl_b2[c] = b2; ctx->cyl_fil[c] = f; ctx->cyl_pos[c] = k;
            ctx->cyl_type[c] = d->fil_type ? d->fil_type[f] : 0;
            ctx->cyl_minus[c] = (k == 0);
            ctx->cyl_kexvol[c] = d->cyl_kexvol[c];
            ctx->cyl_eql[c] = d->cyl_eql[c];
            str_b[2 * c] = b1; str_b[2 * c + 1] = b2;
            str_k[c] = d->cyl_kstr[c]; str_eql[c] = d->cyl_eql[c];
        }
    }
    // FilamentBending::vectorize (FilamentBending.cpp:29-60)
    int n_bend = n_cyl - d->n_filaments;
    std::vector<int> bend_b(3 * (size_t)std::max(0, n_bend));
    std::vector<double> bend_k(std::max(0, n_bend)), bend_eqt(std::max(0, n_bend));
    std::vector<int> cyl_bend(std::max(1, n_cyl), -1);
    ctx->bend_cyl.assign(std::max(0, n_bend), 0);
    {
        int i = 0;
        c = 0;
        for (int f = 0; f < d->n_filaments; f++) {
            int nb = d->fil_bead_offset[f + 1] - d->fil_bead_offset[f];
            for (int k = 1; k < nb - 1; k++, i++) {
                int cp = c + k - 1, cc = c + k;
                bend_b[3 * i] = ctx->cyl_b1[cp];
                bend_b[3 * i + 1] = ctx->cyl_b1[cc];
                bend_b[3 * i + 2] = ctx->cyl_b2[cc];
                bend_k[i] = d->cyl_kbend[cc];
                bend_eqt[i] = d->cyl_eqtheta[cc];
                ctx->bend_cyl[i] = cc;
                cyl_bend[cc] = i;
            }
            c += nb - 1;
        }
    }
    for (int w = 0; w < 2; w++) {
        const int n = w ? d->n_motors : d->n_linkers;
        const mdn_crosslink* xl = w ? d->motors : d->linkers;
        if (n < 0 || (n > 0 && !xl)) return fail(ctx, MDN_ERR_ARG, "bad %s table", w ? "motor" : "linker");
        for (int j = 0; j < n; j++)
            if (xl[j].cyl1 < 0 || xl[j].cyl1 >= n_cyl || xl[j].cyl2 < 0 || xl[j].cyl2 >= n_cyl)
                return fail(ctx, MDN_ERR_ARG, "%s %d: cylinder index out of range", w ? "motor" : "linker", j);
    }
    ctx->elems.assign(d->boundary, d->boundary + d->n_boundary);
    for (auto& e : ctx->elems)
        if (e.kind != MDN_BND_PLANE && e.kind != MDN_BND_SPHERE)
            return fail(ctx, MDN_ERR_ARG, "unknown boundary element kind %d", e.kind);
    ctx->bnd_cutoff = d->boundary_cutoff;
    ctx->rmax = d->exvol_rmax;
    ctx->rmin = d->exvol_rmin;
    for (int q = 0; q < 3; q++) ctx->grid[q] = d->grid_size[q];
    if (!(ctx->rmax > 0)) return fail(ctx, MDN_ERR_ARG, "exvol_rmax must be > 0");

    DevSys& s = ctx->sys;
    s = DevSys{};
    s.n[FF_STRETCH] = n_cyl;
    s.n[FF_BEND] = n_bend;
    s.n[FF_LINKER] = 0;
    s.n[FF_MOTOR] = 0;
    s.n[FF_EXVOL] = 0;
    s.n[FF_BOUNDARY] = 0;
    layout(ctx);

    // per bead, which bonded pieces it has, in the reference's per-bead order: stretching term
    // with the bead as bead 2, as bead 1; bending term as bead 3, 2, 1
    std::vector<unsigned char> pmask(ctx->n_beads, 0);
    {
        auto put = [&](int b, int k) {
            unsigned char& m = pmask[b];
            if (m & (1u << k)) return false;   // a bead listed twice in the filament tables
            m |= (unsigned char)(1u << k);
            return true;
        };
        bool ok = true;
        for (int t = 0; t < n_cyl; t++) ok = put(str_b[2 * t + 1], 0) && put(str_b[2 * t], 1) && ok;
        for (int t = 0; t < n_bend; t++)
            ok = put(bend_b[3 * t + 2], 2) && put(bend_b[3 * t + 1], 3) && put(bend_b[3 * t], 4) && ok;
        if (!ok) return fail(ctx, MDN_ERR_ARG, "a bead appears more than once in the filament bead lists");
    }
    // per bead, its neighbours along the filament and the indices of the stretching term it
    // starts and the bending term it is the middle bead of (cylinders and bending terms are
    // numbered filament after filament, as above)
    const int nbd = ctx->n_beads;
    std::vector<int> bnbr(4 * (size_t)nbd), bterm(2 * (size_t)nbd, 0);
    for (int b = 0; b < nbd; b++)
        for (int q = 0; q < 4; q++) bnbr[(size_t)q * nbd + b] = b;
    for (int f = 0; f < d->n_filaments; f++) {
        const int off = d->fil_bead_offset[f], nf = d->fil_bead_offset[f + 1] - off;
        for (int j = 0; j < nf; j++) {
            const int b = d->fil_beads[off + j];
            const int step[4] = {-2, -1, 1, 2};
            for (int q = 0; q < 4; q++)
                if (j + step[q] >= 0 && j + step[q] < nf) bnbr[(size_t)q * nbd + b] = d->fil_beads[off + j + step[q]];
            bterm[b] = off - f + j;
            bterm[(size_t)nbd + b] = off - 2 * f + j - 1;
        }
    }
    for (int t = 0; t < n_cyl; t++)   // the numbering the recomputation relies on
        if (bterm[str_b[2 * t]] != t || bnbr[2 * (size_t)nbd + str_b[2 * t]] != str_b[2 * t + 1])
            return fail(ctx, MDN_ERR_STATE, "stretching term %d: inconsistent filament tables", t);
    for (int t = 0; t < n_bend; t++)
        if (bterm[(size_t)nbd + bend_b[3 * t + 1]] != t || bnbr[(size_t)nbd + bend_b[3 * t + 1]] != bend_b[3 * t] ||
            bnbr[2 * (size_t)nbd + bend_b[3 * t + 1]] != bend_b[3 * t + 2])
            return fail(ctx, MDN_ERR_STATE, "bending term %d: inconsistent filament tables", t);
    HIPCHK(upload(ctx, ctx->bnbr, bnbr));
    HIPCHK(upload(ctx, ctx->bterm, bterm));
    {   // the bead blocks' neighbours from the indices alone when beads run filament by filament
        const int total = d->n_filaments > 0 ? d->fil_bead_offset[d->n_filaments] : 0;
        bool contig = true;
        for (int k = 0; k < total && contig; k++) contig = d->fil_beads[k] == k;
        const char* env = getenv("MDN_BEAD_TABLES");   // 1: always read the neighbour tables
        ctx->sys.bcontig = contig && !(env && env[0] == '1') ? 1 : 0;
    }
    HIPCHK(ctx->bsum.ensure(3 * (size_t)nbd));
    HIPCHK(upload(ctx, ctx->str_b, str_b)); HIPCHK(upload(ctx, ctx->str_k, str_k)); HIPCHK(upload(ctx, ctx->str_eql, str_eql));
    HIPCHK(upload(ctx, ctx->bend_b, bend_b)); HIPCHK(upload(ctx, ctx->bend_k, bend_k)); HIPCHK(upload(ctx, ctx->bend_eqt, bend_eqt));
    if ((st = set_crosslinks(ctx, d->linkers, d->n_linkers, d->motors, d->n_motors))) return st;
    HIPCHK(upload(ctx, ctx->pmask, pmask));
    HIPCHK(upload(ctx, ctx->d_cyl_bend, cyl_bend));
    HIPCHK(ctx->piece.ensure(NBOND * 3 * (size_t)ctx->n_beads));
    // cylinder tables for the neighbor lists
    HIPCHK(upload(ctx, ctx->d_cyl_b1, ctx->cyl_b1)); HIPCHK(upload(ctx, ctx->d_cyl_b2, ctx->cyl_b2));
    HIPCHK(upload(ctx, ctx->d_cyl_fil, ctx->cyl_fil)); HIPCHK(upload(ctx, ctx->d_cyl_pos, ctx->cyl_pos));
    HIPCHK(upload(ctx, ctx->d_cyl_type, ctx->cyl_type)); HIPCHK(upload(ctx, ctx->d_cyl_minus, ctx->cyl_minus));
    HIPCHK(upload(ctx, ctx->d_cyl_kexvol, ctx->cyl_kexvol)); HIPCHK(upload(ctx, ctx->d_cyl_eql, ctx->cyl_eql));
    {
        std::vector<int> ends(2 * (size_t)std::max(1, ctx->n_fil), 0);
        for (int f = 0, c0 = 0; f < ctx->n_fil; f++) {
            const int nb = d->fil_bead_offset[f + 1] - d->fil_bead_offset[f];
            ends[2 * f] = c0;               // minus-end cylinder
            ends[2 * f + 1] = c0 + nb - 2;  // plus-end cylinder
            c0 += nb - 1;
        }
        HIPCHK(upload(ctx, ctx->d_fil_ends, ends));
    }
    {
        std::vector<double> ke(ctx->elems.size()), se(ctx->elems.size());
        for (size_t e = 0; e < ctx->elems.size(); e++) { ke[e] = ctx->elems[e].krep; se[e] = ctx->elems[e].slen; }
        HIPCHK(upload(ctx, ctx->d_krep_e, ke)); HIPCHK(upload(ctx, ctx->d_slen_e, se));
    }
    // empty dynamic CSR until the neighbor lists are built
    {
        std::vector<int> z(ctx->n_beads + 1, 0);
        HIPCHK(upload(ctx, ctx->inc_d_start, z));
        HIPCHK(ctx->inc_d.ensure(1));
    }
    // state vectors
    size_t nv = 3 * (size_t)ctx->n_beads;
    for (auto* b : {&ctx->x[0], &ctx->x[1], &ctx->f, &ctx->f2, &ctx->fp, &ctx->xh, &ctx->d[0], &ctx->d[1],
                    &ctx->xmin[0], &ctx->xmin[1], &ctx->scratch_f, &ctx->ext})
        HIPCHK(b->ensure(nv));
    ctx->cur = 0;
    if (coord) {
        HIPCHK(hipMemcpyAsync(ctx->x[0].p, coord, sizeof(double) * nv, hipMemcpyHostToDevice, ctx->stream));
    } else {
        HIPCHK(hipMemcpyAsync(ctx->x[0].p, ctx->dof_next.p, sizeof(double) * nv, hipMemcpyDeviceToDevice, ctx->stream));
    }
    ctx->dof_next_n = -1;
    refresh_ptrs(ctx);
    if ((st = ensure_eval_buffers(ctx))) return st;
    refresh_ptrs(ctx);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->uploaded = true;
    return MDN_OK;
}

// serializeDof (DofSerializer.hpp:64-92) from the device coordinates: bead i of the next system is
// current bead src[i], or (src[i] < 0) the next new bead of new_coord.
int mdn_dof_serialize(mdn_ctx* ctx, int32_t n_beads, const int32_t* src, const double* new_coord) {
    int st = need_system(ctx, false);
    if (st) return st;
    if (n_beads <= 0 || !src) return fail(ctx, MDN_ERR_ARG, "n_beads must be positive and src given");
    const int n_old = ctx->n_beads;
    std::vector<char> used(n_old, 0);
    int n_new = 0;
    for (int i = 0; i < n_beads; i++) {
        const int s = src[i];
        if (s < 0) { n_new++; continue; }
        if (s >= n_old) return fail(ctx, MDN_ERR_ARG, "src[%d] = %d: the current system has %d beads", i, s, n_old);
        if (used[s]) return fail(ctx, MDN_ERR_ARG, "src[%d] = %d: current bead named twice", i, s);
        used[s] = 1;
    }
    if (n_new > 0 && !new_coord) return fail(ctx, MDN_ERR_ARG, "%d new beads and no new_coord", n_new);
    HIPCHK(hipSetDevice(ctx->device));
    // the gather map: >= 0 a current bead, -1 - k the k-th new bead
    std::vector<int> g(n_beads);
    for (int i = 0, k = 0; i < n_beads; i++) g[i] = src[i] >= 0 ? src[i] : -1 - k++;
    HIPCHK(upload(ctx, ctx->dof_src, g));
    HIPCHK(ctx->dof_new.ensure(3 * (size_t)std::max(1, n_new)));
    if (n_new > 0)
        HIPCHK(hipMemcpyAsync(ctx->dof_new.p, new_coord, sizeof(double) * 3 * n_new, hipMemcpyHostToDevice,
                              ctx->stream));
    HIPCHK(ctx->dof_next.ensure(3 * (size_t)n_beads));
    const long ne = 3L * n_beads;
    hipLaunchKernelGGL(k_dof_gather, dim3(ceil_div(ne, BLOCK)), dim3(BLOCK), 0, ctx->stream, n_beads,
                       (const int*)ctx->dof_src.p, (const double*)ctx->x[ctx->cur].p, (const double*)ctx->dof_new.p,
                       ctx->dof_next.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));   // g ends with this scope
    ctx->dof_next_n = n_beads;
    return MDN_OK;
}

int mdn_crosslinks_update(mdn_ctx* ctx, const mdn_crosslink* linkers, int32_t n_linkers,
                          const mdn_crosslink* motors, int32_t n_motors) {
    int st = need_system(ctx, false);
    if (st) return st;
    HIPCHK(hipSetDevice(ctx->device));
    for (int w = 0; w < 2; w++) {   // validate everything before any state changes
        const int n = w ? n_motors : n_linkers;
        const mdn_crosslink* xl = w ? motors : linkers;
        if (n < 0 || (n > 0 && !xl)) return fail(ctx, MDN_ERR_ARG, "bad %s table", w ? "motor" : "linker");
        for (int j = 0; j < n; j++)
            if (xl[j].cyl1 < 0 || xl[j].cyl1 >= ctx->n_cyl || xl[j].cyl2 < 0 || xl[j].cyl2 >= ctx->n_cyl)
                return fail(ctx, MDN_ERR_ARG, "%s %d: cylinder index out of range", w ? "motor" : "linker", j);
    }
    if ((st = set_crosslinks(ctx, linkers, n_linkers, motors, n_motors))) return st;
    refresh_ptrs(ctx);
    if (ctx->nl_built) {
        if ((st = relayout(ctx))) return st;
    } else {
        ctx->sys_gen++;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MDN_OK;
}

int mdn_system_counts(mdn_ctx* ctx, int64_t* out) {
    if (!ctx || !out) return MDN_ERR_ARG;
    out[0] = ctx->n_beads;
    out[1] = ctx->n_cyl;
    for (int ff = 0; ff < NFF; ff++) out[2 + ff] = ctx->sys.n[ff];
    return MDN_OK;
}

int mdn_set_coords(mdn_ctx* ctx, const double* coord) {
    int st = need_system(ctx, false);
    if (st) return st;
    if (!coord) return fail(ctx, MDN_ERR_ARG, "null coord");
    st = set_coords_dev(ctx, coord);
    if (st) return st;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MDN_OK;
}

int mdn_displace(mdn_ctx* ctx, const double* delta, int32_t n_var) {
    int st = need_system(ctx, false);
    if (st) return st;
    if (!delta || n_var != 3 * ctx->n_beads) return fail(ctx, MDN_ERR_ARG, "n_var must be 3*n_beads = %d", 3 * ctx->n_beads);
    HIPCHK(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_axpy_inplace, dim3(std::max(1, ceil_div(n_var, BLOCK))), dim3(BLOCK), 0, ctx->stream,
                       ctx->x[ctx->cur].p, delta, (long)n_var);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MDN_OK;
}

int mdn_perturb(mdn_ctx* ctx, uint64_t seed, uint64_t stream, double sigma) {
    int st = need_system(ctx, false);
    if (st) return st;
    if (!(sigma >= 0.0)) return fail(ctx, MDN_ERR_ARG, "sigma must be >= 0");
    HIPCHK(hipSetDevice(ctx->device));
    const long n = 3L * ctx->n_beads;
    hipLaunchKernelGGL(k_perturb, dim3(std::max(1, ceil_div(n, BLOCK))), dim3(BLOCK), 0, ctx->stream,
                       ctx->x[ctx->cur].p, n, (unsigned long long)seed, (unsigned long long)stream, sigma);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MDN_OK;
}

int mdn_get_coords(mdn_ctx* ctx, double* coord) {
    int st = need_system(ctx, false);
    if (st) return st;
    HIPCHK(hipMemcpyAsync(coord, ctx->x[ctx->cur].p, sizeof(double) * 3 * ctx->n_beads, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MDN_OK;
}

// The block / slot layout for the current interaction counts, the dynamic incidence (excluded-
// volume and boundary slots per bead, from the tables of the last neighbor-list build) and the
// merged bead-sorted slot layout with the static crosslink incidence.
int relayout(mdn_ctx* ctx) {
    int st = MDN_OK;
    const long np = ctx->sys.n[FF_EXVOL];
    const int nbi = ctx->sys.n[FF_BOUNDARY];
    layout(ctx);
    if ((st = ensure_eval_buffers(ctx))) return st;
    refresh_ptrs(ctx);
    // dynamic incidence CSR
    HIPCHK(ctx->inc_cnt.ensure(ctx->n_beads + 1));
    HIPCHK(ctx->inc_cursor.ensure(ctx->n_beads + 1));
    HIPCHK(ctx->inc_d_start.ensure(ctx->n_beads + 1));
    const long ndyn = 4L * np + nbi;
    HIPCHK(ctx->inc_d.ensure(std::max(1L, ndyn)));
    HIPCHK(hipMemsetAsync(ctx->inc_cnt.p, 0, sizeof(int) * (ctx->n_beads + 1), ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->inc_cursor.p, 0, sizeof(int) * (ctx->n_beads + 1), ctx->stream));
    refresh_ptrs(ctx);
    if (ndyn > 0)
        hipLaunchKernelGGL(k_inc_count, dim3(ceil_div(ndyn, BLOCK)), dim3(BLOCK), 0, ctx->stream, ctx->sys,
                           ctx->inc_cnt.p);
    if ((st = exclusive_scan(ctx, ctx->inc_cnt.p, ctx->inc_d_start.p, ctx->n_beads + 1))) return st;
    if (ndyn > 0) {
        hipLaunchKernelGGL(k_inc_fill, dim3(ceil_div(ndyn, BLOCK)), dim3(BLOCK), 0, ctx->stream, ctx->sys,
                           ctx->inc_d_start.p, ctx->inc_cursor.p, ctx->inc_d.p);
        hipLaunchKernelGGL(k_inc_sort, dim3(bead_blocks(ctx)), dim3(BLOCK), 0, ctx->stream, ctx->n_beads,
                           ctx->inc_d_start.p, ctx->inc_d.p);
    }
    // bead-sorted slot layout (static then dynamic slots of each bead)
    HIPCHK(ctx->bstart.ensure(ctx->n_beads + 1));
    hipLaunchKernelGGL(k_inc_total, dim3(bead_blocks(ctx)), dim3(BLOCK), 0, ctx->stream, ctx->n_beads,
                       ctx->inc_s_start.p, ctx->inc_d_start.p, ctx->inc_cnt.p);
    HIPCHK(hipMemsetAsync(ctx->inc_cnt.p + ctx->n_beads, 0, sizeof(int), ctx->stream));
    if ((st = exclusive_scan(ctx, ctx->inc_cnt.p, ctx->bstart.p, ctx->n_beads + 1))) return st;
    hipLaunchKernelGGL(k_inc_merge, dim3(bead_blocks(ctx)), dim3(BLOCK), 0, ctx->stream, ctx->n_beads,
                       ctx->inc_s_start.p, ctx->inc_s.p, ctx->inc_d_start.p, ctx->inc_d.p, ctx->bstart.p,
                       ctx->sorted_id.p, ctx->pos.p);
    HIPCHK(hipGetLastError());
    refresh_ptrs(ctx);
    ctx->sys_gen++;
    return MDN_OK;
}

// The buffers of one hybrid neighbor list build (midpoints, bins, members, per-cylinder counts
// and offsets, the pair array).
struct NlBufs {
    DBuf<double>& mid;
    DBuf<int>& bin; DBuf<int>& bin_count; DBuf<int>& bin_start; DBuf<int>& bin_cursor; DBuf<int>& members;
    DBuf<int>& count; DBuf<int>& off; DBuf<int>& pairs;
    DBuf<int>& mbin; DBuf<double>& mmid;   // each member's bin and midpoint, in bin order
};

// HybridCylinderCylinderNL::reset for one list: midpoints, counting sort into bins, the
// 27-bin stencil (count, scan, emit, per-cylinder sort). *np = the pairs, ordered by cylinder
// then neighbour.
static int build_pairs(mdn_ctx* ctx, const NLDev& n, const double* x, NlBufs& B, long* np, const char* what) {
    int st = MDN_OK;
    const long ntiles = (long)n.nt[0] * n.nt[1] * n.nt[2];
    const long nbins = ntiles * n.T * n.T * n.T;   // tile by tile, the last tiles padded
    if (nbins > (1L << 30)) return fail(ctx, MDN_ERR_UNSUPPORTED, "neighbor-list grid too large (%ld bins)", nbins);
    const int nc = n.n_cyl;
    HIPCHK(B.mid.ensure(3 * (size_t)std::max(1, nc)));
    HIPCHK(B.bin.ensure(std::max(1, nc)));
    HIPCHK(B.bin_count.ensure(nbins + 1));
    HIPCHK(B.bin_start.ensure(nbins + 1));
    HIPCHK(B.bin_cursor.ensure(nbins + 1));
    HIPCHK(B.members.ensure(std::max(1, nc)));
    HIPCHK(B.mbin.ensure(std::max(1, nc)));
    HIPCHK(B.mmid.ensure(3 * (size_t)std::max(1, nc)));
    HIPCHK(B.count.ensure(nc + 1));
    HIPCHK(B.off.ensure(nc + 1));
    HIPCHK(hipMemsetAsync(B.bin_count.p, 0, sizeof(int) * (nbins + 1), ctx->stream));
    HIPCHK(hipMemsetAsync(B.bin_cursor.p, 0, sizeof(int) * (nbins + 1), ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->nl_err.p, 0, sizeof(int), ctx->stream));
    const int G = std::max(1, ceil_div(nc, BLOCK));
    hipLaunchKernelGGL(k_nl_bin, dim3(G), dim3(BLOCK), 0, ctx->stream, n, x, B.mid.p, B.bin.p, B.bin_count.p,
                       ctx->nl_err.p);
    HIPCHK(hipGetLastError());
    if ((st = exclusive_scan(ctx, B.bin_count.p, B.bin_start.p, nbins + 1))) return st;
    hipLaunchKernelGGL(k_nl_fill, dim3(G), dim3(BLOCK), 0, ctx->stream, nc, B.bin.p, B.bin_start.p, B.bin_cursor.p,
                       (const double*)B.mid.p, B.members.p, B.mbin.p, B.mmid.p);
    hipLaunchKernelGGL(k_nl_tile<0>, dim3(ntiles), dim3(BLOCK), 0, ctx->stream, n, (const int*)B.bin_start.p,
                       (const int*)B.members.p, (const int*)B.mbin.p, (const double*)B.mmid.p, B.count.p,
                       (const int*)nullptr, (int*)nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemsetAsync(B.count.p + nc, 0, sizeof(int), ctx->stream));
    if ((st = exclusive_scan(ctx, B.count.p, B.off.p, nc + 1))) return st;
    HIPCHK(hipMemcpyAsync(ctx->h_int, B.off.p + nc, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->h_int + 1, ctx->nl_err.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->h_int[1] != 0) {
        return fail(ctx, MDN_ERR_NL_BOUNDS, "%d cylinder midpoints outside the neighbor-list grid [0, %g)x[0, %g)x[0, %g)",
                    ctx->h_int[1], ctx->grid[0], ctx->grid[1], ctx->grid[2]);
    }
    *np = ctx->h_int[0];
    if (*np > 0x7fffffffL / 4) return fail(ctx, MDN_ERR_UNSUPPORTED, "too many %s pairs (%ld)", what, *np);
    HIPCHK(B.pairs.ensure(2 * (size_t)std::max(1L, *np)));
    hipLaunchKernelGGL(k_nl_tile<1>, dim3(ntiles), dim3(BLOCK), 0, ctx->stream, n, (const int*)B.bin_start.p,
                       (const int*)B.members.p, (const int*)B.mbin.p, (const double*)B.mmid.p, B.count.p,
                       (const int*)B.off.p, B.pairs.p);
    HIPCHK(hipGetLastError());
    return MDN_OK;
}

int mdn_nl_build(mdn_ctx* ctx, const double* coord, int64_t* n_pairs_out) {
    int st = need_system(ctx, false);
    if (st) return st;
    HIPCHK(hipSetDevice(ctx->device));
    // The rebuild reallocates the pair, excluded-volume and boundary tables that the system
    // record and the captured minimizer graphs point at; until the new layout is complete
    // (relayout), nothing may use them: a failure on the way leaves the lists unbuilt.
    ctx->nl_built = false;
    if ((st = set_coords_dev(ctx, coord))) return st;
    const double* x = ctx->x[ctx->cur].p;
    NLDev n = nl_view(ctx);
    const int nc = ctx->n_cyl;
    long np = 0;
    {
        NlBufs B{ctx->mid, ctx->bin, ctx->bin_count, ctx->bin_start, ctx->bin_cursor, ctx->members,
                 ctx->nl_count, ctx->nl_off, ctx->pairs, ctx->mbin, ctx->mmid};
        if ((st = build_pairs(ctx, n, x, B, &np, "excluded-volume"))) return st;
    }
    // CylinderExclVolume::vectorize
    HIPCHK(ctx->ev_b.ensure(4 * (size_t)std::max(1L, np)));
    HIPCHK(ctx->ev_ks.ensure((size_t)std::max(1L, np)));
    if (np > 0)
        hipLaunchKernelGGL(k_ev_vectorize, dim3(ceil_div(np, BLOCK)), dim3(BLOCK), 0, ctx->stream, n, np,
                           ctx->pairs.p, ctx->ev_b.p, ctx->ev_ks.p);
    ctx->n_pairs = np;
    // boundary NL + BoundaryCylinderRepulsion::vectorize
    const int ne = (int)ctx->elems.size();
    const long nbc = (long)ne * nc;
    ctx->sys.n_elem = ne;
    HIPCHK(ctx->bnd_cnt.ensure(nbc + 1));
    HIPCHK(ctx->bnd_off.ensure(nbc + 1));
    if (nbc > 0)
        hipLaunchKernelGGL(k_bnd_count, dim3(ceil_div(nbc, BLOCK)), dim3(BLOCK), 0, ctx->stream, n, ctx->sys,
                           ctx->bnd_cutoff, ctx->mid.p, ctx->bnd_cnt.p);
    HIPCHK(hipMemsetAsync(ctx->bnd_cnt.p + nbc, 0, sizeof(int), ctx->stream));
    if ((st = exclusive_scan(ctx, ctx->bnd_cnt.p, ctx->bnd_off.p, nbc + 1))) return st;
    HIPCHK(hipMemcpyAsync(ctx->h_int, ctx->bnd_off.p + nbc, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    // per-element counts (nneighbors) come from the offsets at element boundaries
    std::vector<int> eoff(ne + 1, 0);
    for (int e = 0; e <= ne; e++)
        HIPCHK(hipMemcpyAsync(&eoff[e], ctx->bnd_off.p + (long)e * nc, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int nbi = ctx->h_int[0];
    ctx->bnd_nneighbors.assign(ne, 0);
    for (int e = 0; e < ne; e++) ctx->bnd_nneighbors[e] = eoff[e + 1] - eoff[e];
    HIPCHK(ctx->bnd_b.ensure(std::max(1, nbi)));
    HIPCHK(ctx->bnd_e.ensure(std::max(1, nbi)));
    HIPCHK(ctx->bnd_krep.ensure(std::max(1, nbi)));
    HIPCHK(ctx->bnd_slen.ensure(std::max(1, nbi)));
    if (nbc > 0)
        hipLaunchKernelGGL(k_bnd_emit, dim3(ceil_div(nbc, BLOCK)), dim3(BLOCK), 0, ctx->stream, n, ctx->sys,
                           ctx->bnd_cnt.p, ctx->bnd_off.p, ctx->d_krep_e.p, ctx->d_slen_e.p, ctx->bnd_b.p,
                           ctx->bnd_e.p, ctx->bnd_krep.p, ctx->bnd_slen.p);
    ctx->n_bint = nbi;
    // new layout
    ctx->sys.n[FF_EXVOL] = (int)np;
    ctx->sys.n[FF_BOUNDARY] = nbi;
    if ((st = relayout(ctx))) return st;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->nl_built = true;
    if (n_pairs_out) *n_pairs_out = np;
    return MDN_OK;
}

int mdn_boundary_load_forces(mdn_ctx* ctx, const double* coord, int32_t n_types, const double* monomer_size,
                             const int32_t* num_monomers, int32_t stride, double* load_plus, double* load_minus) {
    int st = need_system(ctx, true);
    if (st) return st;
    if (!monomer_size || !num_monomers || !load_plus || !load_minus) return fail(ctx, MDN_ERR_ARG, "null argument");
    if (n_types < 1 || n_types > MAX_FIL_TYPES)
        return fail(ctx, MDN_ERR_ARG, "n_types must be in [1, %d]", MAX_FIL_TYPES);
    LoadParams P;
    memset(&P, 0, sizeof(P));
    P.n_types = n_types;
    P.stride = stride;
    for (int t = 0; t < n_types; t++) {
        if (num_monomers[t] < 0 || num_monomers[t] > stride)
            return fail(ctx, MDN_ERR_ARG, "num_monomers[%d] = %d outside [0, stride = %d]", t, num_monomers[t], stride);
        P.mon_size[t] = monomer_size[t];
        P.num_mon[t] = num_monomers[t];
    }
    for (int c = 0; c < ctx->n_cyl; c++)
        if (ctx->cyl_type[c] < 0 || ctx->cyl_type[c] >= n_types)
            return fail(ctx, MDN_ERR_ARG, "filament %d has type %d, not below n_types = %d", ctx->cyl_fil[c],
                        ctx->cyl_type[c], n_types);
    HIPCHK(hipSetDevice(ctx->device));
    const double* x = ctx->x[ctx->cur].p;
    if (coord) {
        HIPCHK(hipMemcpyAsync(ctx->ext.p, coord, sizeof(double) * 3 * ctx->n_beads, hipMemcpyHostToDevice, ctx->stream));
        x = ctx->ext.p;
    }
    const long n_out = 2L * ctx->n_fil * stride;
    if (n_out > 0) {
        HIPCHK(ctx->load_out.ensure(n_out));
        hipLaunchKernelGGL(k_load_forces, dim3(ceil_div(n_out, BLOCK)), dim3(BLOCK), 0, ctx->stream, nl_view(ctx),
                           ctx->sys, x, (const int*)ctx->d_fil_ends.p, (const int*)ctx->bnd_cnt.p,
                           (const double*)ctx->d_krep_e.p, (const double*)ctx->d_slen_e.p, ctx->n_fil, P,
                           ctx->load_out.p);
        HIPCHK(hipGetLastError());
        const size_t half = sizeof(double) * (size_t)ctx->n_fil * stride;
        HIPCHK(hipMemcpyAsync(load_plus, ctx->load_out.p, half, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(load_minus, ctx->load_out.p + (size_t)ctx->n_fil * stride, half, hipMemcpyDeviceToHost,
                              ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MDN_OK;
}

// HybridBindingSearchManager::updateAllPossibleBindingsstencilHYBD (HybridBindingSearchManager.cpp:
// 759-966): the binder's hybrid list (addtoHNeighborList, :1024-1039: range cyl_size + rmax,
// lower bound max(rmin - cyl_size, 0), in float), then the site pairs of every listed pair,
// counted, scanned and emitted in list order.
int mdn_binding_pairs(mdn_ctx* ctx, const mdn_binding_params* p, const double* coord, uint32_t* out,
                      int64_t capacity, int64_t* n_out) {
    int st = need_system(ctx, false);
    if (st) return st;
    if (!p || !n_out) return fail(ctx, MDN_ERR_ARG, "null argument");
    if (p->n_sites <= 0 || !p->site_pos) return fail(ctx, MDN_ERR_ARG, "n_sites must be positive, site_pos given");
    if (p->shift_bits <= 0 || p->shift_bits > 16 || p->n_sites > (1 << p->shift_bits))
        return fail(ctx, MDN_ERR_ARG, "shift_bits %d cannot hold %d sites", p->shift_bits, p->n_sites);
    if (((long)ctx->n_cyl << p->shift_bits) > 0xffffffffL)
        return fail(ctx, MDN_ERR_UNSUPPORTED, "%d cylinders << %d bits exceed 32 bits", ctx->n_cyl, p->shift_bits);
    if (!(p->rmax > 0) || p->rmin < 0 || p->site_step < 0) return fail(ctx, MDN_ERR_ARG, "bad range or site step");
    HIPCHK(hipSetDevice(ctx->device));
    const double* x = ctx->x[ctx->cur].p;
    if (coord) {
        HIPCHK(hipMemcpyAsync(ctx->ext.p, coord, sizeof(double) * 3 * ctx->n_beads, hipMemcpyHostToDevice, ctx->stream));
        x = ctx->ext.p;
    }
    // the list's range in float, as addtoHNeighborList and setneighborsearchparameters form it
    const float rmaxsq_b = p->rmax * p->rmax, rminsq_b = p->rmin * p->rmin;
    const float rmax_l = p->cyl_size + std::sqrt(rmaxsq_b);
    const float rmin_l = std::max(std::sqrt(rminsq_b) - p->cyl_size, 0.0f);
    NLDev n = nl_view(ctx);
    n.rmaxsq = rmax_l * rmax_l;
    n.rminsq = rmin_l * rmin_l;
    n.side = 1.125 * std::sqrt(n.rmaxsq);
    for (int q = 0; q < 3; q++) n.nb[q] = std::max(1, (int)std::ceil(ctx->grid[q] / n.side));
    nl_tiles(n);
    n.ta = std::min(p->ftype1, p->ftype2);
    n.tb = std::max(p->ftype1, p->ftype2);
    long np = 0;
    {
        NlBufs B{ctx->b_mid, ctx->b_bin, ctx->b_bin_count, ctx->b_bin_start, ctx->b_bin_cursor, ctx->b_members,
                 ctx->b_count, ctx->b_off, ctx->b_pairs, ctx->b_mbin, ctx->b_mmid};
        if ((st = build_pairs(ctx, n, x, B, &np, "binding-list"))) return st;
    }
    BindDev b{};
    b.n_sites = p->n_sites;
    b.shift = p->shift_bits;
    b.step = p->site_step;
    b.ftype1 = std::min(p->ftype1, p->ftype2);   // the reference's _filamentIDvec pair, ascending
    b.ftype2 = std::max(p->ftype1, p->ftype2);
    b.rminsq = rminsq_b;
    b.rmaxsq = rmaxsq_b;
    b.p2_first = p->p2_first;
    b.p2_last = p->p2_last;
    {
        std::vector<float> site(p->site_pos, p->site_pos + p->n_sites);
        HIPCHK(upload(ctx, ctx->b_site, site));
        if (p->site_free) {
            std::vector<float> fr(p->site_free, p->site_free + (size_t)ctx->n_cyl * p->n_sites);
            HIPCHK(upload(ctx, ctx->b_free, fr));
        }
        HIPCHK(hipStreamSynchronize(ctx->stream));   // the host vectors end with this scope
    }
    b.site = ctx->b_site.p;
    b.free = p->site_free ? ctx->b_free.p : nullptr;
    HIPCHK(ctx->b_scount.ensure(np + 1));
    HIPCHK(ctx->b_soff.ensure(np + 1));
    const int G = std::max(1, ceil_div(np, BLOCK));
    hipLaunchKernelGGL(k_bind_sites<0>, dim3(G), dim3(BLOCK), 0, ctx->stream, n, b, x, np, (const int*)ctx->b_pairs.p,
                       ctx->b_scount.p, (const int*)nullptr, (uint32_t*)nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemsetAsync(ctx->b_scount.p + np, 0, sizeof(int), ctx->stream));
    if ((st = exclusive_scan(ctx, ctx->b_scount.p, ctx->b_soff.p, np + 1))) return st;
    HIPCHK(d2h(ctx, ctx->h_int, ctx->b_soff.p + np, sizeof(int)));
    const long total = ctx->h_int[0];
    *n_out = total;
    if (!out) return MDN_OK;
    if (capacity < total)
        return fail(ctx, MDN_ERR_CAPACITY, "%ld binding pairs, capacity %ld", total, (long)capacity);
    if (total == 0) return MDN_OK;
    HIPCHK(ctx->b_out.ensure(2 * (size_t)total));
    hipLaunchKernelGGL(k_bind_sites<1>, dim3(G), dim3(BLOCK), 0, ctx->stream, n, b, x, np, (const int*)ctx->b_pairs.p,
                       ctx->b_scount.p, (const int*)ctx->b_soff.p, ctx->b_out.p);
    HIPCHK(hipGetLastError());
    HIPCHK(d2h(ctx, out, ctx->b_out.p, sizeof(uint32_t) * 2 * total));
    return MDN_OK;
}

int mdn_nl_pairs(mdn_ctx* ctx, int32_t* out, int64_t cap, int64_t* n_pairs) {
    int st = need_system(ctx, true);
    if (st) return st;
    if (n_pairs) *n_pairs = ctx->n_pairs;
    if (cap < ctx->n_pairs) return fail(ctx, MDN_ERR_CAPACITY, "need room for %ld pairs", ctx->n_pairs);
    if (ctx->n_pairs > 0)
        HIPCHK(d2h(ctx, out, ctx->pairs.p, sizeof(int) * 2 * ctx->n_pairs));
    return MDN_OK;
}

int mdn_boundary_list(mdn_ctx* ctx, int32_t* out_beads, int64_t cap, int32_t* out_nneighbors,
                      int64_t* n_interactions) {
    int st = need_system(ctx, true);
    if (st) return st;
    if (n_interactions) *n_interactions = ctx->n_bint;
    if (cap < ctx->n_bint) return fail(ctx, MDN_ERR_CAPACITY, "need room for %d interactions", ctx->n_bint);
    if (ctx->n_bint > 0 && out_beads)
        HIPCHK(d2h(ctx, out_beads, ctx->bnd_b.p, sizeof(int) * ctx->n_bint));
    if (out_nneighbors)
        for (size_t e = 0; e < ctx->bnd_nneighbors.size(); e++) out_nneighbors[e] = ctx->bnd_nneighbors[e];
    return MDN_OK;
}

static int energy_impl(mdn_ctx* ctx, const double* coord, mdn_energy_report* out, int prec) {
    int st = need_system(ctx, true);
    if (st) return st;
    if (!out) return fail(ctx, MDN_ERR_ARG, "null report");
    HIPCHK(hipSetDevice(ctx->device));
    const double* x = ctx->x[ctx->cur].p;
    if (coord) {
        HIPCHK(hipMemcpyAsync(ctx->ext.p, coord, sizeof(double) * 3 * ctx->n_beads, hipMemcpyHostToDevice, ctx->stream));
        x = ctx->ext.p;
    }
    if ((st = energy_report(ctx, x, prec))) return st;
    fill_report(ctx->h_eout, out);
    return MDN_OK;
}

int mdn_energy(mdn_ctx* ctx, const double* coord, mdn_energy_report* out) {
    return energy_impl(ctx, coord, out, MDN_PREC_FP64);
}

int mdn_energy_mixed(mdn_ctx* ctx, const double* coord, mdn_energy_report* out) {
    return energy_impl(ctx, coord, out, MDN_PREC_MIXED);
}

static int forces_impl(mdn_ctx* ctx, int ff, const double* coord, double* force, int32_t n_var, int prec) {
    int st = need_system(ctx, true);
    if (st) return st;
    if (!force || n_var != 3 * ctx->n_beads) return fail(ctx, MDN_ERR_ARG, "n_var must be 3*n_beads = %d", 3 * ctx->n_beads);
    HIPCHK(hipSetDevice(ctx->device));
    const double* x = ctx->x[ctx->cur].p;
    if (coord) {
        HIPCHK(hipMemcpyAsync(ctx->ext.p, coord, sizeof(double) * n_var, hipMemcpyHostToDevice, ctx->stream));
        x = ctx->ext.p;
    }
    long lo = 0, hi = ctx->sys.slot0[NFF];
    const double* base = nullptr;
    if (ff >= 0) {
        lo = ctx->sys.slot0[ff];
        hi = ctx->sys.slot0[ff + 1];
        HIPCHK(hipMemcpyAsync(ctx->scratch_f.p, force, sizeof(double) * n_var, hipMemcpyHostToDevice, ctx->stream));
        base = ctx->scratch_f.p;
    }
    const int bmask = ff < 0 ? 3 : ff == FF_STRETCH ? 1 : ff == FF_BEND ? 2 : 0;
    const bool bg = ctx->bg_eval != 0;
    if ((st = eval_slots(ctx, x, prec, !bg, bmask, base))) return st;
    if ((st = gather(ctx, bmask, lo, hi, base, ctx->f.p, prec, bg))) return st;
    HIPCHK(hipMemcpyAsync(force, ctx->f.p, sizeof(double) * n_var, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MDN_OK;
}

int mdn_forces(mdn_ctx* ctx, const double* coord, double* force, int32_t n_var) {
    return forces_impl(ctx, -1, coord, force, n_var, MDN_PREC_FP64);
}

int mdn_ff_forces(mdn_ctx* ctx, int32_t ff, const double* coord, double* force, int32_t n_var) {
    if (ff < 0 || ff >= NFF) return fail(ctx, MDN_ERR_ARG, "force field index %d", ff);
    return forces_impl(ctx, ff, coord, force, n_var, MDN_PREC_FP64);
}

int mdn_crosslink_stretch_forces(mdn_ctx* ctx, double* linker_out, double* motor_out) {
    int st = need_system(ctx, true);
    if (st) return st;
    // the tensions are stored in the device tables' order: back to the crosslinks' own order
    for (int w = 0; w < 2; w++) {
        double* out = w ? motor_out : linker_out;
        const int n = ctx->sys.n[w ? FF_MOTOR : FF_LINKER];
        if (!out || n == 0) continue;
        std::vector<double> tmp(n);
        HIPCHK(d2h(ctx, tmp.data(), (w ? ctx->mot_sf : ctx->lnk_sf).p, sizeof(double) * n));
        const std::vector<int>& ord = ctx->xl_ord[w];
        for (int i = 0; i < n; i++) out[ord[i]] = tmp[i];
    }
    return MDN_OK;
}

// dE/dtheta at device coordinates x into ctx->grad_out: the per-pair excluded-volume energies,
// then one thread per cylinder. check: settle the pairs whose closed-form energy is deferred (a
// host sync on a device count, then the Simpson energy of each listed pair); without it the
// caller has established that no pair at x is deferred.
static int param_grad_launch(mdn_ctx* ctx, const double* x, int prec, bool check) {
    const DevSys& s = ctx->sys;
    const int np = s.n[FF_EXVOL];
    if (np > 0) {
        HIPCHK(hipMemsetAsync(ctx->bad_count.p, 0, sizeof(int), ctx->stream));
        hipLaunchKernelGGL(k_ev_pair_energy, dim3(ceil_div(np, BLOCK)), dim3(BLOCK), 0, ctx->stream, s, x, ctx->ev_U.p,
                           ctx->bad_count.p, ctx->bad_list.p);
        HIPCHK(hipGetLastError());
    }
    if (np > 0 && check) {
        HIPCHK(hipMemcpyAsync(ctx->h_int, ctx->bad_count.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        const int nbad = ctx->h_int[0];
        if (nbad > 0) {
            std::vector<int> redo(nbad);
            HIPCHK(d2h(ctx, redo.data(), ctx->bad_list.p, sizeof(int) * nbad));
            std::sort(redo.begin(), redo.end());
            for (int t : redo)
                hipLaunchKernelGGL(k_ev_pair_energyN, dim3(1), dim3(BLOCK), 0, ctx->stream, s, x, t, ctx->ev_U.p);
            HIPCHK(hipGetLastError());
            std::vector<double> U(nbad);
            for (int i = 0; i < nbad; i++)
                HIPCHK(hipMemcpyAsync(&U[i], ctx->ev_U.p + redo[i], sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            for (int i = 0; i < nbad; i++)
                if (U[i] <= -1.0)
                    return fail(ctx, MDN_ERR_NONFINITE,
                                "excluded-volume pair %d: energy non-finite after the Simpson fallback", redo[i]);
        }
    }
    if (ctx->n_cyl > 0) {
        if (prec == MDN_PREC_MIXED)
            hipLaunchKernelGGL(k_param_grad<float>, dim3(ceil_div(ctx->n_cyl * PG_LANES, BLOCK)), dim3(BLOCK), 0, ctx->stream, s,
                               ctx->n_cyl, (const int*)ctx->d_cyl_b1.p, (const int*)ctx->d_cyl_bend.p, x,
                               (const double*)ctx->ev_U.p, ctx->grad_out.p);
        else
            hipLaunchKernelGGL(k_param_grad<double>, dim3(ceil_div(ctx->n_cyl * PG_LANES, BLOCK)), dim3(BLOCK), 0, ctx->stream, s,
                               ctx->n_cyl, (const int*)ctx->d_cyl_b1.p, (const int*)ctx->d_cyl_bend.p, x,
                               (const double*)ctx->ev_U.p, ctx->grad_out.p);
    }
    HIPCHK(hipGetLastError());
    return MDN_OK;
}

static int param_grad_impl(mdn_ctx* ctx, const double* coord, double* out, int64_t n_out, int prec) {
    int st = need_system(ctx, true);
    if (st) return st;
    if (!out || n_out != 3L * ctx->n_cyl) return fail(ctx, MDN_ERR_ARG, "n_out must be 3*n_cylinders = %d", 3 * ctx->n_cyl);
    HIPCHK(hipSetDevice(ctx->device));
    const double* x = ctx->x[ctx->cur].p;
    if (coord) {
        HIPCHK(hipMemcpyAsync(ctx->ext.p, coord, sizeof(double) * 3 * ctx->n_beads, hipMemcpyHostToDevice, ctx->stream));
        x = ctx->ext.p;
    }
    HIPCHK(ctx->ev_U.ensure(std::max(1, ctx->sys.n[FF_EXVOL])));
    HIPCHK(ctx->grad_out.ensure(3 * (size_t)std::max(1, ctx->n_cyl)));
    if ((st = param_grad_launch(ctx, x, prec, true))) return st;
    HIPCHK(hipMemcpyAsync(out, ctx->grad_out.p, sizeof(double) * 3 * ctx->n_cyl, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MDN_OK;
}

int mdn_param_grad(mdn_ctx* ctx, const double* coord, double* out, int64_t n_out) {
    return param_grad_impl(ctx, coord, out, n_out, MDN_PREC_FP64);
}

int mdn_param_grad_mixed(mdn_ctx* ctx, const double* coord, double* out, int64_t n_out) {
    return param_grad_impl(ctx, coord, out, n_out, MDN_PREC_MIXED);
}

// The tangent terms of nv products at device coordinates x: product k's tangent at v + k n_var,
// its slots at slots + k slot_stride, its pieces at piece + k piece_stride. Excluded-volume pairs
// without a tangent are counted in bad_count (zeroed by the caller).
static void launch_tangents(mdn_ctx* ctx, const double* x, int prec, const double* v, int nv, double* slots,
                            long slot_stride, double* piece, long piece_stride) {
    const DevSys& s = ctx->sys;
    if (s.blk[NFF] == 0 || nv == 0) return;
    const long n_var = 3L * ctx->n_beads;
    const dim3 g(s.blk[NFF], nv);
    if (prec == MDN_PREC_MIXED)
        hipLaunchKernelGGL(k_tangent_terms<DualF>, g, dim3(BLOCK), 0, ctx->stream, s, x, v, n_var, slots, slot_stride,
                           piece, piece_stride, ctx->bad_count.p);
    else
        hipLaunchKernelGGL(k_tangent_terms<DualD>, g, dim3(BLOCK), 0, ctx->stream, s, x, v, n_var, slots, slot_stride,
                           piece, piece_stride, ctx->bad_count.p);
}

// (dF/dx) v of nv products into out + k n_var from their slots and pieces (one launch).
static void gather_tangents(mdn_ctx* ctx, int nv, const double* slots, long slot_stride, double* piece,
                            long piece_stride, double* out) {
    const DevSys& s = ctx->sys;
    const dim3 g(gather_blocks(ctx), nv), b(BLOCK);
    const long n_var = 3L * ctx->n_beads;
    if (gather_size(ctx) == 128)
        hipLaunchKernelGGL(k_gather_tangents<128>, g, b, 0, ctx->stream, s, slots, slot_stride, piece, piece_stride, out,
                           n_var);
    else
        hipLaunchKernelGGL(k_gather_tangents<64>, g, b, 0, ctx->stream, s, slots, slot_stride, piece, piece_stride, out,
                           n_var);
}

// (dF/dx) v of one product into out from its slots and pieces (the FP64 gather).
static void gather_tangent(mdn_ctx* ctx, const double* slots, double* piece, double* out) {
    DevSys s = ctx->sys;
    s.piece = piece;
    launch_gather_plain<double>(ctx, s, 3, slots, 0L, s.slot0[NFF], (const double*)nullptr, out, false);
}

// -H v into ctx->scratch_f for the device tangent ctx->tan_v at device coordinates x (the tangent
// pieces of every interaction, then the FP64 gather); excluded-volume pairs without a tangent are
// counted in bad_count.
static int hvp_launch(mdn_ctx* ctx, const double* x, int prec) {
    HIPCHK(hipMemsetAsync(ctx->bad_count.p, 0, sizeof(int), ctx->stream));
    launch_tangents(ctx, x, prec, ctx->tan_v.p, 1, ctx->slots.p, 0L, ctx->piece.p, 0L);
    HIPCHK(hipGetLastError());
    gather_tangent(ctx, ctx->slots.p, ctx->piece.p, ctx->scratch_f.p);
    HIPCHK(hipGetLastError());
    return MDN_OK;
}

static int hvp_impl(mdn_ctx* ctx, const double* coord, const double* v, double* out, int32_t n_var, int prec) {
    int st = need_system(ctx, true);
    if (st) return st;
    if (!v || !out || n_var != 3 * ctx->n_beads) return fail(ctx, MDN_ERR_ARG, "n_var must be 3*n_beads = %d", 3 * ctx->n_beads);
    HIPCHK(hipSetDevice(ctx->device));
    const double* x = ctx->x[ctx->cur].p;
    if (coord) {
        HIPCHK(hipMemcpyAsync(ctx->ext.p, coord, sizeof(double) * n_var, hipMemcpyHostToDevice, ctx->stream));
        x = ctx->ext.p;
    }
    HIPCHK(ctx->tan_v.ensure(n_var));
    HIPCHK(hipMemcpyAsync(ctx->tan_v.p, v, sizeof(double) * n_var, hipMemcpyHostToDevice, ctx->stream));
    if ((st = hvp_launch(ctx, x, prec))) return st;
    HIPCHK(hipMemcpyAsync(out, ctx->scratch_f.p, sizeof(double) * n_var, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->h_int, ctx->bad_count.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->h_int[0] != 0)
        return fail(ctx, MDN_ERR_UNSUPPORTED, "%d excluded-volume pairs need the numerical force fallback (no tangent)",
                    ctx->h_int[0]);
    for (int i = 0; i < n_var; i++) out[i] = -out[i];   // H v = -(dF/dx) v
    return MDN_OK;
}

int mdn_hvp(mdn_ctx* ctx, const double* coord, const double* v, double* out, int32_t n_var) {
    return hvp_impl(ctx, coord, v, out, n_var, MDN_PREC_FP64);
}

int mdn_hvp_mixed(mdn_ctx* ctx, const double* coord, const double* v, double* out, int32_t n_var) {
    return hvp_impl(ctx, coord, v, out, n_var, MDN_PREC_MIXED);
}

int mdn_hvp_batch(mdn_ctx* ctx, int32_t precision, const double* coord, const double* V, double* out, int32_t n_var,
                  int32_t n_vec) {
    int st = need_system(ctx, true);
    if (st) return st;
    if (precision != MDN_PREC_FP64 && precision != MDN_PREC_MIXED)
        return fail(ctx, MDN_ERR_ARG, "unknown precision %d", precision);
    if (!V || !out || n_var != 3 * ctx->n_beads) return fail(ctx, MDN_ERR_ARG, "n_var must be 3*n_beads = %d", 3 * ctx->n_beads);
    if (n_vec < 1) return fail(ctx, MDN_ERR_ARG, "n_vec must be positive");
    HIPCHK(hipSetDevice(ctx->device));
    const DevSys& s = ctx->sys;
    const double* x = ctx->x[ctx->cur].p;
    if (coord) {
        HIPCHK(hipMemcpyAsync(ctx->ext.p, coord, sizeof(double) * n_var, hipMemcpyHostToDevice, ctx->stream));
        x = ctx->ext.p;
    }
    // per product: its tangent, its slots (with the gather's padding), its pieces, its result
    const int chunk = std::min<int>(n_vec, MDN_HVP_BATCH_CHUNK);
    const long slot_stride = 3 * std::max(1L, s.slot0[NFF]) + GCHUNK_PAD;
    const long piece_stride = (long)NBOND * 3 * ctx->n_beads;
    HIPCHK(ctx->hb_v.ensure((size_t)chunk * n_var));
    HIPCHK(ctx->hb_slots.ensure((size_t)chunk * slot_stride));
    HIPCHK(ctx->hb_piece.ensure((size_t)chunk * piece_stride));
    HIPCHK(ctx->hb_out.ensure((size_t)chunk * n_var));
    HIPCHK(hipMemsetAsync(ctx->bad_count.p, 0, sizeof(int), ctx->stream));
    for (int k0 = 0; k0 < n_vec; k0 += chunk) {
        const int nv = std::min(chunk, n_vec - k0);
        HIPCHK(hipMemcpyAsync(ctx->hb_v.p, V + (size_t)k0 * n_var, sizeof(double) * nv * n_var, hipMemcpyHostToDevice,
                              ctx->stream));
        launch_tangents(ctx, x, precision, ctx->hb_v.p, nv, ctx->hb_slots.p, slot_stride, ctx->hb_piece.p, piece_stride);
        HIPCHK(hipGetLastError());
        gather_tangents(ctx, nv, ctx->hb_slots.p, slot_stride, ctx->hb_piece.p, piece_stride, ctx->hb_out.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out + (size_t)k0 * n_var, ctx->hb_out.p, sizeof(double) * nv * n_var,
                              hipMemcpyDeviceToHost, ctx->stream));
    }
    HIPCHK(hipMemcpyAsync(ctx->h_int, ctx->bad_count.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->h_int[0] != 0)
        return fail(ctx, MDN_ERR_UNSUPPORTED, "%d excluded-volume pairs need the numerical force fallback (no tangent)",
                    ctx->h_int[0]);
    const size_t total = (size_t)n_vec * n_var;
    for (size_t i = 0; i < total; i++) out[i] = -out[i];   // H v = -(dF/dx) v
    return MDN_OK;
}

int mdn_set_params(mdn_ctx* ctx, const double* kstr, const double* eql, const double* kbend) {
    int st = need_system(ctx, false);
    if (st) return st;
    HIPCHK(hipSetDevice(ctx->device));
    const int nc = ctx->n_cyl;
    if (kstr) HIPCHK(hipMemcpyAsync(ctx->str_k.p, kstr, sizeof(double) * nc, hipMemcpyHostToDevice, ctx->stream));
    if (eql) {
        for (int c = 0; c < nc; c++) ctx->cyl_eql[c] = eql[c];
        HIPCHK(upload(ctx, ctx->str_eql, ctx->cyl_eql));
        HIPCHK(upload(ctx, ctx->d_cyl_eql, ctx->cyl_eql));
        if (ctx->nl_built && ctx->n_pairs > 0)   // kRepScaled = krep L0_i L0_j
            hipLaunchKernelGGL(k_ev_vectorize, dim3(ceil_div(ctx->n_pairs, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                               nl_view(ctx), ctx->n_pairs, (const int*)ctx->pairs.p, ctx->ev_b.p, ctx->ev_ks.p);
    }
    if (kbend) {
        std::vector<double> bk(ctx->bend_cyl.size());
        for (size_t i = 0; i < bk.size(); i++) bk[i] = kbend[ctx->bend_cyl[i]];
        HIPCHK(upload(ctx, ctx->bend_k, bk));
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MDN_OK;
}

int mdn_forces_mixed(mdn_ctx* ctx, const double* coord, double* force, int32_t n_var) {
    return forces_impl(ctx, -1, coord, force, n_var, MDN_PREC_MIXED);
}

}  // extern "C"

// ----------------------------------------------------------------------------------------
// Force evaluation without allocation
// ----------------------------------------------------------------------------------------
namespace {

// Byte offsets of the evaluation's buffers in `scratch` (each 256-byte aligned): the fallback
// count and flag, the fallback partials (6 per block of the 240-block Simpson rule), the list of
// pairs that need it, the bonded sums (3 doubles per bead) and last the slot vectors (3 values of
// R per slot, padded by one gather chunk, which the gather reads unconditionally).
struct ScratchLayout {
    size_t bad_count, ok, partial, bad_list, bsum, slots, total;
};
constexpr int FORCEN_BLOCKS = 240;
size_t up256(size_t v) { return (v + 255) & ~(size_t)255; }
ScratchLayout scratch_layout(const DevSys& s, int n_beads, size_t rsize) {
    ScratchLayout L;
    L.bad_count = 0;
    L.ok = 256;
    L.partial = 512;
    L.bad_list = up256(L.partial + sizeof(double) * 6 * FORCEN_BLOCKS);
    L.bsum = up256(L.bad_list + sizeof(int) * (size_t)std::max(1, s.n[FF_EXVOL]));
    L.slots = up256(L.bsum + sizeof(double) * 3 * (size_t)std::max(1, n_beads));
    L.total = L.slots + rsize * (3 * (size_t)std::max(1L, s.slot0[NFF]) + GCHUNK_PAD);
    return L;
}

template <typename R>
int forces_noalloc_t(mdn_ctx* ctx, unsigned char* scr, const double* x, double* out) {
    DevSys s = ctx->sys;   // the same tables, with the bonded sums in the scratch
    const ScratchLayout L = scratch_layout(s, ctx->n_beads, sizeof(R));
    int* bad_count = reinterpret_cast<int*>(scr + L.bad_count);
    int* ok = reinterpret_cast<int*>(scr + L.ok);
    double* partial = reinterpret_cast<double*>(scr + L.partial);
    int* bad_list = reinterpret_cast<int*>(scr + L.bad_list);
    s.bsum = reinterpret_cast<double*>(scr + L.bsum);
    R* slots = reinterpret_cast<R*>(scr + L.slots);
    // the previous call's count stays in the array: clear it (a kernel, not a memset node)
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, ctx->stream, bad_count, 0);
    hipLaunchKernelGGL(k_forces_terms<R>, dim3(force_blocks(s, false)), dim3(BLOCK), 0, ctx->stream, s, x, slots,
                       bad_count, bad_list, bead_block0(s, false), 3, (const double*)nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(d2h(ctx, ctx->h_int, bad_count, sizeof(int)));
    const int nbad = ctx->h_int[0];
    // pairs whose closed-form forces are non-finite: the Simpson forceN rule, retried once with
    // beads 1 and 2 moved out of plane (CylinderExclVolRepulsion.cpp:1198-1204); each pair writes
    // only its own slots, so the order does not matter
    for (int i = 0; i < nbad; i++) {
        HIPCHK(d2h(ctx, ctx->h_int + 1, bad_list + i, sizeof(int)));
        const int t = ctx->h_int[1];
        int done = 0;
        for (int mv = 0; mv < 2 && !done; mv++) {
            hipLaunchKernelGGL(k_exvol_forceN_rows, dim3(FORCEN_BLOCKS), dim3(BLOCK), 0, ctx->stream, s, x, t, mv,
                               partial);
            hipLaunchKernelGGL(k_exvol_forceN_finish<R>, dim3(1), dim3(64), 0, ctx->stream, s, x, t, mv,
                               FORCEN_BLOCKS, (const double*)partial, slots, ok);
            HIPCHK(hipGetLastError());
            HIPCHK(d2h(ctx, ctx->h_int + 2, ok, sizeof(int)));
            done = ctx->h_int[2];
        }
        if (!done)
            return fail(ctx, MDN_ERR_NONFINITE,
                        "excluded-volume pair %d: force non-finite after the numerical fallback", t);
    }
    launch_gather_plain<R>(ctx, s, 3, (const R*)slots, 0L, s.slot0[NFF], (const double*)nullptr, out, true);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MDN_OK;
}

}  // namespace

extern "C" {

int mdn_forces_noalloc(mdn_ctx* ctx, int32_t precision, const double* coord_dev, double* force_dev, int32_t n_var) {
    int st = need_system(ctx, true);
    if (st) return st;
    if (!coord_dev || !force_dev) return fail(ctx, MDN_ERR_ARG, "null coordinate or force buffer");
    if (n_var != 3 * ctx->n_beads) return fail(ctx, MDN_ERR_ARG, "n_var must be 3*n_beads = %d", 3 * ctx->n_beads);
    if (precision != MDN_PREC_FP64 && precision != MDN_PREC_MIXED)
        return fail(ctx, MDN_ERR_ARG, "unknown precision %d", precision);
    const ScratchLayout L =
        scratch_layout(ctx->sys, ctx->n_beads, precision == MDN_PREC_MIXED ? sizeof(float) : sizeof(double));
    if (L.total > MDN_SCRATCH_BYTES)
        return fail(ctx, MDN_ERR_CAPACITY,
                    "the evaluation needs %zu bytes of scratch; mdn_forces_noalloc has %zu (%d MiB): use mdn_forces",
                    L.total, MDN_SCRATCH_BYTES, MDN_NOALLOC_SCRATCH_MIB);
    HIPCHK(hipSetDevice(ctx->device));
    unsigned char* scr = ctx->device < MAX_DEVICES ? g_scratch_dev[ctx->device] : nullptr;
    if (!scr) return fail(ctx, MDN_ERR_STATE, "no scratch address for device %d", ctx->device);
    return precision == MDN_PREC_MIXED ? forces_noalloc_t<float>(ctx, scr, coord_dev, force_dev)
                                       : forces_noalloc_t<double>(ctx, scr, coord_dev, force_dev);
}

// ----------------------------------------------------------------------------------------
// CGMethod::minimize on the device
// ----------------------------------------------------------------------------------------
// Line-search partial blocks: the interaction blocks (k_cg_ls), then the bead blocks (k_cg_pos).
static int ls_blocks(mdn_ctx* ctx) { return ctx->sys.lsp; }

static CgBufs cg_bufs(mdn_ctx* ctx) {
    CgBufs B;
    B.x[0] = ctx->x[0].p; B.x[1] = ctx->x[1].p;
    B.d[0] = ctx->d[0].p; B.d[1] = ctx->d[1].p;
    B.xmin[0] = ctx->xmin[0].p; B.xmin[1] = ctx->xmin[1].p;
    B.f[0] = ctx->f.p; B.f[1] = ctx->f2.p; B.fp = ctx->fp.p; B.xh = ctx->xh.p;
    B.slots = (void*)ctx->slots.p; B.partial = ctx->partial.p; B.bad_count = ctx->bad_count.p;
    B.gpartial = ctx->partial.p + (long)PSTRIDE * ls_blocks(ctx);   // after the line-search partials
    return B;
}

// The minimizer's kernels use the bonded sums (bonded_gather); below CG_BEAD_ENERGY_BEADS beads
// the line search keeps the filament terms' own blocks for their energies, from there on the bead
// blocks take the energies too and those blocks leave the grid (bead_energy). Measured per CG
// iteration (same process, piece arrays -> this choice): 42,000 beads FP64 32.2 -> 31.2-32.3 us,
// mixed 30.5 -> 29.6 us (bead-block energies there: 32.9 and 30.8 us); 420,000 beads FP64 75.3 ->
// 68.4 us, mixed 65.5 -> 63.0 us (own energy blocks there: 73.1 and 67.9 us).
static const int CG_BEAD_ENERGY_BEADS = 100000;
static bool cg_bonded_sums(const mdn_ctx* ctx, int prec) {
    (void)prec;
    return ctx->bonded_gather_mode >= 0 ? ctx->bonded_gather_mode != 0 : true;
}
// the system record the minimizer's kernels of precision prec take
static DevSys cg_sys(const mdn_ctx* ctx, int prec) {
    DevSys s = ctx->sys;
    s.bonded_gather = cg_bonded_sums(ctx, prec) ? 1 : 0;
    s.bead_energy = ctx->n_beads >= CG_BEAD_ENERGY_BEADS ? 1 : 0;
    return s;
}

extern "C++" {
template <typename R>
static void launch_slot(mdn_ctx* ctx, const DevSys& s, const CgBufs& B) {
    const int gbk = gather_blocks(ctx);
    hipLaunchKernelGGL(k_cg_pos, dim3(bead_blocks(ctx)), dim3(BLOCK), 0, ctx->stream, s, (const CgState*)ctx->cg.p, B);
    // all interaction blocks, then (bonded_gather) the bead blocks; with bead_energy the filament
    // terms' blocks are left out
    const int lsg = s.bonded_gather ? (s.bead_energy ? s.blk_nb : s.blk[NFF]) + s.nbb : s.blk[NFF];
    hipLaunchKernelGGL(k_cg_ls<R>, dim3(std::max(1, lsg)), dim3(BLOCK), 0, ctx->stream, s,
                       (const CgState*)ctx->cg.p, B);
    launch_cg_gather<R>(ctx, s, (const CgState*)ctx->cg.p, B);
    hipLaunchKernelGGL(k_cg_fin, dim3(1), dim3(FIN), 0, ctx->stream, s, ctx->cg.p, B, gbk);
}
}  // extern "C++"

// A batch of minimizer slots captured once per system layout and replayed as a hipGraph; the
// kernels take every decision from the device state, so replays need no host work.
static const int CG_BATCH = 16;

static int cg_graph(mdn_ctx* ctx, int prec) {
    if (ctx->graph_exec[prec] && ctx->graph_gen[prec] == ctx->sys_gen) return MDN_OK;
    if (ctx->graph_exec[prec]) (void)hipGraphExecDestroy(ctx->graph_exec[prec]);
    if (ctx->graph[prec]) (void)hipGraphDestroy(ctx->graph[prec]);
    ctx->graph_exec[prec] = nullptr;
    ctx->graph[prec] = nullptr;
    CgBufs B = cg_bufs(ctx);
    const DevSys s = cg_sys(ctx, prec);
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < CG_BATCH; i++) {
        if (prec == MDN_PREC_MIXED) launch_slot<float>(ctx, s, B);
        else launch_slot<double>(ctx, s, B);
    }
    HIPCHK(hipStreamEndCapture(ctx->stream, &ctx->graph[prec]));
    HIPCHK(hipGraphInstantiate(&ctx->graph_exec[prec], ctx->graph[prec], nullptr, nullptr, 0));
    ctx->graph_gen[prec] = ctx->sys_gen;
    return MDN_OK;
}

int mdn_minimize(mdn_ctx* ctx, const mdn_cg_params* p, double* coord, double* force_out, mdn_min_result* res) {
    int st = need_system(ctx, true);
    if (st) return st;
    if (!p || !res) return fail(ctx, MDN_ERR_ARG, "null argument");
    if (p->precision != MDN_PREC_FP64 && p->precision != MDN_PREC_MIXED)
        return fail(ctx, MDN_ERR_ARG, "unknown precision %d", p->precision);
    const int prec = p->precision;
    if (p->method < 0 || p->method > 2) return fail(ctx, MDN_ERR_ARG, "unknown descent method %d", p->method);
    if (p->line_search != MDN_LS_BACKTRACKING && p->line_search != MDN_LS_QUADRATIC)
        return fail(ctx, MDN_ERR_ARG, "unknown line search %d", p->line_search);
    HIPCHK(hipSetDevice(ctx->device));
    memset(res, 0, sizeof(*res));
    const int nv = 3 * ctx->n_beads;
    // startMinimization + initial energy / forces (CGMethod.hpp:876-896); the start is in x[0]
    if (coord)
        HIPCHK(hipMemcpyAsync(ctx->x[0].p, coord, sizeof(double) * nv, hipMemcpyHostToDevice, ctx->stream));
    else if (ctx->cur != 0)   // the context's device coordinates
        HIPCHK(hipMemcpyAsync(ctx->x[0].p, ctx->x[ctx->cur].p, sizeof(double) * nv, hipMemcpyDeviceToDevice,
                              ctx->stream));
    ctx->cur = 0;
    // the reported energies are FP64 in both precisions; the minimizer's own energies and forces
    // are evaluated in the requested precision
    if ((st = energy_report(ctx, ctx->x[0].p, MDN_PREC_FP64))) return st;
    fill_report(ctx->h_eout, &res->energies_before);
    double e_start = res->energies_before.total;
    if (prec == MDN_PREC_MIXED) {
        if ((st = energy_report(ctx, ctx->x[0].p, prec))) return st;
        e_start = ctx->h_eout[6];
    }
    if ((st = eval_slots(ctx, ctx->x[0].p, prec, !ctx->bg_eval))) return st;
    if ((st = gather(ctx, 3, 0, ctx->sys.slot0[NFF], nullptr, ctx->f.p, prec, ctx->bg_eval != 0))) return st;
    // F.F and maxF on the device (fixed-order reduction), two doubles back
    const int nsb = std::min(VSTAT_BLOCKS, std::max(1, ceil_div(nv, BLOCK)));
    hipLaunchKernelGGL(k_vec_stats, dim3(nsb), dim3(BLOCK), 0, ctx->stream, (const double*)ctx->f.p, (long)nv,
                       ctx->vstat.p);
    hipLaunchKernelGGL(k_vec_stats_fin, dim3(1), dim3(BLOCK), 0, ctx->stream, (const double*)ctx->vstat.p, nsb,
                       ctx->eout.p + 16);
    HIPCHK(hipMemcpyAsync(ctx->h_eout + 16, ctx->eout.p + 16, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const double ff = ctx->h_eout[16], mx = ctx->h_eout[17];
    CgState& S = *ctx->h_cg;
    memset(&S, 0, sizeof(S));
    S.method = p->method;
    S.ls_algo = p->line_search;
    S.try_recover = p->try_recover;
    S.force_call_limit = std::max(p->force_call_limit, 3 * ctx->n_beads);
    S.energy_call_limit = p->energy_call_limit;
    S.max_dist = p->max_dist;
    S.lambda_max = p->lambda_max;
    S.grad_tol = p->grad_tol;
    S.energy_rel_tol = p->energy_rel_tol;
    S.num_energy_call = 1;
    S.num_force_call = 1;
    S.grad_inf_norm = mx;
    S.energy_rel_change = INFINITY;
    S.prev_energy = e_start;
    S.minimum_e = INFINITY;
    S.fdotf = ff;
    // resetCGParams before the loop (d = F + 0 d, lastFDotF = F.F, F_prev = F), applied by the
    // first line-search launch
    S.dir_pending = 1;
    S.dir_beta = 0.0;
    S.dir_reset = 1;
    S.phase = PH_LS;
    S.ls_lambda = rmin(S.lambda_max, S.max_dist / S.grad_inf_norm);
    res->force_call_limit = S.force_call_limit;
    res->energy_call_limit = S.energy_call_limit;
    res->grad_tol = S.grad_tol;
    res->energy_rel_tol = S.energy_rel_tol;
    if (!std::isfinite(mx) || !std::isfinite(ff)) return fail(ctx, MDN_ERR_NONFINITE, "maxF is not finite at the start");
    const bool limit0 = (S.force_call_limit != 0 && S.num_force_call > S.force_call_limit);
    S.done = (S.grad_inf_norm < S.grad_tol || S.energy_rel_change < S.energy_rel_tol || limit0) ? 1 : 0;
    HIPCHK(hipMemsetAsync(ctx->d[0].p, 0, sizeof(double) * nv, ctx->stream));   // searchDir.assign(numDof, 0)
    HIPCHK(hipMemcpyAsync(ctx->cg.p, &S, sizeof(CgState), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->bad_count.p, 0, sizeof(int), ctx->stream));
    // The slots are captured into a graph once a layout is minimized a second time: a graph made
    // for one minimization (a layout that the next neighbor-list rebuild replaces) costs more to
    // instantiate than its replays save over launching the same kernels directly.
    const bool graph = ctx->use_graph && (ctx->graph_gen[prec] == ctx->sys_gen || ctx->min_gen[prec] == ctx->sys_gen);
    ctx->min_gen[prec] = ctx->sys_gen;
    if (graph && (st = cg_graph(ctx, prec))) return st;
    const DevSys slot_sys = cg_sys(ctx, prec);
    const CgBufs slot_bufs = cg_bufs(ctx);

    // Batches are replayed back to back; the host polls the done flag of batch i while batch
    // i+1 runs, so the device never idles waiting for the host.
    const long max_batches = (slots_per_force_call(S) * ((long)S.force_call_limit + 16)) / CG_BATCH + 8;
    long launched = 0;
    hipEvent_t ev[2];
    HIPCHK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    int n_forcen = 0;
    for (;;) {
        bool done = S.done != 0;
        int* h_done = ctx->h_int + 4;
        h_done[0] = h_done[1] = 0;
        int k = 0;
        long here = 0;
        while (!done && launched < max_batches) {
            if (graph) {
                HIPCHK(hipGraphLaunch(ctx->graph_exec[prec], ctx->stream));
            } else {
                for (int i = 0; i < CG_BATCH; i++) {
                    if (prec == MDN_PREC_MIXED) launch_slot<float>(ctx, slot_sys, slot_bufs);
                    else launch_slot<double>(ctx, slot_sys, slot_bufs);
                }
                HIPCHK(hipGetLastError());
            }
            launched++;
            here++;
            HIPCHK(hipMemcpyAsync(&h_done[k], &ctx->cg.p->done, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipEventRecord(ev[k], ctx->stream));
            if (here > 1) {
                HIPCHK(hipEventSynchronize(ev[k ^ 1]));
                if (h_done[k ^ 1]) done = true;
            }
            k ^= 1;
        }
        HIPCHK(d2h(ctx, &S, ctx->cg.p, sizeof(CgState)));
        if (!(S.done && S.status == ST_NEED_FORCEN)) break;
        // The step's forces need the numerical excluded-volume fallback (forceN,
        // CylinderExclVolRepulsion.cpp:984-1254): recompute them at the new coordinates with it
        // (same positions, so every other slot is bit-identical), then run this step's gather
        // again and resume.
        n_forcen++;
        const DevSys cs = cg_sys(ctx, prec);
        if ((st = eval_slots(ctx, ctx->x[S.parity ^ 1].p, prec, !cs.bonded_gather))) return st;
        HIPCHK(hipMemsetAsync(ctx->bad_count.p, 0, sizeof(int), ctx->stream));
        S.status = 0;
        S.done = 0;
        S.reserved[1] = n_forcen;
        HIPCHK(hipMemcpyAsync(ctx->cg.p, &S, sizeof(CgState), hipMemcpyHostToDevice, ctx->stream));
        const CgBufs B = cg_bufs(ctx);
        const int gbk = gather_blocks(ctx);
        if (prec == MDN_PREC_MIXED) launch_cg_gather<float>(ctx, cs, (const CgState*)ctx->cg.p, B);
        else launch_cg_gather<double>(ctx, cs, (const CgState*)ctx->cg.p, B);
        hipLaunchKernelGGL(k_cg_fin, dim3(1), dim3(FIN), 0, ctx->stream, cs, ctx->cg.p, B, gbk);
        HIPCHK(hipGetLastError());
        // the state after this step's finalize (and the host state's upload above) is complete
        // before the pinned copy is overwritten
        HIPCHK(d2h(ctx, &S, ctx->cg.p, sizeof(CgState)));
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    ctx->last_cg = S;
    if (!S.done) return fail(ctx, MDN_ERR_STATE, "minimizer did not terminate within %ld batches", launched);
    ctx->cur = S.parity;
    res->error_flags = S.error_flags;
    res->num_energy_call = S.num_energy_call;
    res->num_force_call = S.num_force_call;
    res->num_iter = S.num_iter;
    res->num_safe_mode = S.num_safe_mode;
    res->num_recovery = S.num_recovery;
    res->grad_inf_norm = S.grad_inf_norm;
    res->energy_rel_change = S.energy_rel_change;
    if (S.status == 9)
        return fail(ctx, MDN_ERR_UNSUPPORTED, "more excluded-volume pairs need the Simpson energy than the "
                    "minimizer's capacity (%d per line-search launch)", DEFER_CAP);
    if (S.status != 0) return fail(ctx, S.status, "minimization stopped with status %d", S.status);
    const bool limit = (S.energy_call_limit != 0 && S.num_energy_call > S.energy_call_limit) ||
                       (S.force_call_limit != 0 && S.num_force_call > S.force_call_limit);
    res->call_limit_exceeded = limit;
    res->converged_grad = S.grad_inf_norm < S.grad_tol;
    res->converged_energy = S.energy_rel_change < S.energy_rel_tol;
    double* xres = ctx->x[ctx->cur].p;
    if (S.error_flags == 0) {
        if (p->report_on_call_limit && limit && S.have_min_e) {   // copybackupcoordinates
            HIPCHK(hipMemcpyAsync(xres, ctx->xmin[S.bkpar].p, sizeof(double) * nv, hipMemcpyDeviceToDevice,
                                  ctx->stream));
            res->restored_min_energy = 1;
        }
        if ((st = energy_report(ctx, xres, MDN_PREC_FP64))) return st;
        fill_report(ctx->h_eout, &res->energies_after);
        res->num_energy_call += 1;
    }
    if (coord) HIPCHK(hipMemcpyAsync(coord, xres, sizeof(double) * nv, hipMemcpyDeviceToHost, ctx->stream));
    if (force_out)
        HIPCHK(hipMemcpyAsync(force_out, (S.fpar ? ctx->f2 : ctx->f).p, sizeof(double) * nv, hipMemcpyDeviceToHost,
                              ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (S.error_flags & MDN_MIN_ERR_LINESEARCH) return fail(ctx, MDN_ERR_LINESEARCH, "line search failed (recovery disabled)");
    return MDN_OK;
}

}  // extern "C"

// ----------------------------------------------------------------------------------------
// Batched minimizer: many independent systems, one persistent kernel
// ----------------------------------------------------------------------------------------
static void fill_result(const CgState& S, const double* rep_before, const double* rep_after, bool have_after,
                        mdn_min_result* res) {
    memset(res, 0, sizeof(*res));
    fill_report(rep_before, &res->energies_before);
    if (have_after) fill_report(rep_after, &res->energies_after);
    res->error_flags = S.error_flags;
    res->num_energy_call = S.num_energy_call + (have_after ? 1 : 0);
    res->num_force_call = S.num_force_call;
    res->force_call_limit = S.force_call_limit;
    res->energy_call_limit = S.energy_call_limit;
    res->num_iter = S.num_iter;
    res->num_safe_mode = S.num_safe_mode;
    res->num_recovery = S.num_recovery;
    res->grad_inf_norm = S.grad_inf_norm;
    res->grad_tol = S.grad_tol;
    res->energy_rel_change = S.energy_rel_change;
    res->energy_rel_tol = S.energy_rel_tol;
    res->call_limit_exceeded = call_limit_host(S);
    res->converged_grad = S.grad_inf_norm < S.grad_tol;
    res->converged_energy = S.energy_rel_change < S.energy_rel_tol;
    res->restored_min_energy = S.reserved[0];
}

template <typename R, int NT>
static int launch_batched(mdn_ctx* L, int n, const BatchParams& P) {
    auto kern = k_cg_batched<R, NT>;
    int per_cu = 0, n_cu = 0;
    HIPCHK_CTX(L, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, NT, 0));
    HIPCHK_CTX(L, hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, L->device));
    int grid = std::max(1, std::min(n, std::max(1, per_cu) * std::max(1, n_cu)));
    // MDN_BATCH_MAX_WORKGROUPS caps the persistent grid, so that every workgroup pulls several
    // systems from the queue (the schedule a batch larger than the resident workgroups runs)
    if (const char* env = getenv("MDN_BATCH_MAX_WORKGROUPS"))
        if (atoi(env) > 0) grid = std::min(grid, atoi(env));
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, L->stream, L->batch_queue.p, 0);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), 0, L->stream, (const BatchItem*)L->batch_items.p,
                       (const int*)L->batch_order.p, n, L->batch_queue.p, P);
    HIPCHK_CTX(L, hipGetLastError());
    return MDN_OK;
}

template <typename R>
static int run_batched_t(mdn_ctx* L, int n, const BatchParams& P, int nt) {
    switch (nt) {
    case 128: return launch_batched<R, 128>(L, n, P);
    case 512: return launch_batched<R, 512>(L, n, P);
    case 1024: return launch_batched<R, 1024>(L, n, P);
    default: return launch_batched<R, 256>(L, n, P);
    }
}

static int run_batched(mdn_ctx* L, int n, const BatchParams& P, int prec, int nt) {
    return prec == MDN_PREC_MIXED ? run_batched_t<float>(L, n, P, nt) : run_batched_t<double>(L, n, P, nt);
}

extern "C" {

int mdn_minimize_batched(mdn_ctx* const* ctxs, int32_t n, const mdn_cg_params* p, double* const* coords,
                         double* const* forces_out, mdn_min_result* results) {
    if (!ctxs || n <= 0 || !ctxs[0]) return MDN_ERR_ARG;
    mdn_ctx* L = ctxs[0];
    if (!p || !coords || !results) return fail(L, MDN_ERR_ARG, "null argument");
    if (p->precision != MDN_PREC_FP64 && p->precision != MDN_PREC_MIXED)
        return fail(L, MDN_ERR_ARG, "unknown precision %d", p->precision);
    if (p->method < 0 || p->method > 2) return fail(L, MDN_ERR_ARG, "unknown descent method %d", p->method);
    if (p->line_search != MDN_LS_BACKTRACKING && p->line_search != MDN_LS_QUADRATIC)
        return fail(L, MDN_ERR_ARG, "unknown line search %d", p->line_search);
    long beads = 0;
    for (int i = 0; i < n; i++) {
        mdn_ctx* c = ctxs[i];
        if (!c) return fail(L, MDN_ERR_ARG, "system %d: null context", i);
        if (c->device != L->device) return fail(L, MDN_ERR_ARG, "system %d is on another device", i);
        for (int j = 0; j < i; j++)
            if (ctxs[j] == c) return fail(L, MDN_ERR_ARG, "system %d repeats context %d", i, j);
        int st = need_system(c, true);
        if (st) return fail(L, st, "system %d: %s", i, c->err.c_str());
        beads += c->n_beads;
    }
    HIPCHK_CTX(L, hipSetDevice(L->device));
    // every context's own work is complete before the lead stream reads its buffers
    for (int i = 1; i < n; i++) HIPCHK_CTX(L, hipStreamSynchronize(ctxs[i]->stream));
    std::vector<BatchItem> items(n);
    std::vector<long> size(n);
    // device-to-device copies in three sets of n (a zero count copies nothing), each set one launch:
    // the systems that start from their device coordinates (x[cur] -> x[0]), the start's backup for
    // the schedule tuning runs (x[0] -> ext) and its restore before each of them (ext -> x[0])
    std::vector<BatchCopy> copies(3 * (size_t)n);
    long max_nv = 1;
    for (int i = 0; i < n; i++) {
        mdn_ctx* c = ctxs[i];
        const long nv = 3L * c->n_beads;
        const size_t bytes = sizeof(double) * nv;
        max_nv = std::max(max_nv, nv);
        if (coords[i])
            HIPCHK_CTX(L, hipMemcpyAsync(c->x[0].p, coords[i], bytes, hipMemcpyHostToDevice, L->stream));
        else if (c->cur != 0)   // the context's device coordinates
            copies[i] = BatchCopy{c->x[c->cur].p, c->x[0].p, nv};
        c->cur = 0;
        copies[n + i] = BatchCopy{c->x[0].p, c->ext.p, nv};
        copies[2L * n + i] = BatchCopy{c->ext.p, c->x[0].p, nv};
        items[i].s = c->sys;
        items[i].B = cg_bufs(c);
        items[i].st = c->cg.p;
        items[i].report = c->eout.p;
        size[i] = c->n_beads;
        for (int ff = 0; ff < NFF; ff++) size[i] += c->sys.n[ff];
    }
    BatchParams P{};
    P.method = p->method;
    P.line_search = p->line_search;
    P.try_recover = p->try_recover;
    P.force_call_limit = p->force_call_limit;
    P.energy_call_limit = p->energy_call_limit;
    P.report_on_call_limit = p->report_on_call_limit;
    P.limit_override = 0;
    P.max_dist = p->max_dist;
    P.lambda_max = p->lambda_max;
    P.grad_tol = p->grad_tol;
    P.energy_rel_tol = p->energy_rel_tol;
    HIPCHK_CTX(L, L->batch_items.ensure(n));
    HIPCHK_CTX(L, L->batch_order.ensure(n));
    HIPCHK_CTX(L, L->batch_queue.ensure(1));
    HIPCHK_CTX(L, L->batch_copy.ensure(3 * (size_t)n));
    HIPCHK_CTX(L, hipMemcpyAsync(L->batch_items.p, items.data(), sizeof(BatchItem) * n, hipMemcpyHostToDevice,
                                 L->stream));
    HIPCHK_CTX(L, hipMemcpyAsync(L->batch_copy.p, copies.data(), sizeof(BatchCopy) * 3 * n, hipMemcpyHostToDevice,
                                 L->stream));
    const dim3 cgrid((unsigned)std::min(64L, std::max(1L, (long)ceil_div(max_nv, 256L))), (unsigned)std::min(n, 4096));
    auto copy_set = [&](int set) {   // set 0: start copies, 1: backups, 2: restores
        hipLaunchKernelGGL(k_batch_copy, cgrid, dim3(256), 0, L->stream,
                           (const BatchCopy*)(L->batch_copy.p + (size_t)set * n), n);
        return hipGetLastError();
    };
    HIPCHK_CTX(L, copy_set(0));
    std::vector<int> order_given(n), order_sorted(n);
    for (int i = 0; i < n; i++) order_given[i] = order_sorted[i] = i;
    std::stable_sort(order_sorted.begin(), order_sorted.end(), [&](int a, int b) { return size[a] > size[b]; });
    int st;
    // schedule: workgroup size (128, 256, 512, 1024) x system order (given, largest first), chosen once
    // per (systems, precision, beads) by timing each candidate on this batch, run to completion
    // from the caller's coordinates: the batch time is set by its slowest systems, whose iteration
    // counts only a complete run shows (a short run measures the start-up instead)
    const long key = ((long)n * 2 + p->precision) * 1000003L + beads;
    if (L->batch_key != key) {
        const BatchParams T = P;
        hipEvent_t e0, e1;
        HIPCHK_CTX(L, hipEventCreate(&e0));
        HIPCHK_CTX(L, hipEventCreate(&e1));
        float best = 1e30f;
        const int nts[4] = {128, 256, 512, 1024};
        HIPCHK_CTX(L, copy_set(1));   // the start, kept for the tuning runs
        for (int c = 0; c < MDN_BATCH_CANDIDATES; c++) {
            const int nt = nts[c >> 1], sorted = c & 1;
            HIPCHK_CTX(L, copy_set(2));   // every candidate starts from the given coordinates
            HIPCHK_CTX(L, hipMemcpyAsync(L->batch_order.p, sorted ? order_sorted.data() : order_given.data(),
                                         sizeof(int) * n, hipMemcpyHostToDevice, L->stream));
            HIPCHK_CTX(L, hipEventRecord(e0, L->stream));
            if ((st = run_batched(L, n, T, p->precision, nt))) return st;
            HIPCHK_CTX(L, hipEventRecord(e1, L->stream));
            HIPCHK_CTX(L, hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK_CTX(L, hipEventElapsedTime(&ms, e0, e1));
            L->batch_tune_ms[c] = ms;
            if (ms < best) {
                best = ms;
                L->batch_nt = nt;
                L->batch_sorted = sorted;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        L->batch_key = key;
        L->batch_tuned = 1;
        // the tuning runs moved the coordinates: start again from the given ones
        HIPCHK_CTX(L, copy_set(2));
    } else {
        L->batch_tuned = 0;
    }
    HIPCHK_CTX(L, hipMemcpyAsync(L->batch_order.p, L->batch_sorted ? order_sorted.data() : order_given.data(),
                                 sizeof(int) * n, hipMemcpyHostToDevice, L->stream));
    if ((st = run_batched(L, n, P, p->precision, L->batch_nt))) return st;
    HIPCHK_CTX(L, hipStreamSynchronize(L->stream));
    // results: every system's state and report gathered on the device and read back at once, then
    // the coordinates (and forces) of every system queued with one synchronization
    HIPCHK_CTX(L, L->batch_states.ensure((size_t)n));
    HIPCHK_CTX(L, L->batch_reports.ensure(27 * (size_t)n));
    hipLaunchKernelGGL(k_batch_collect, dim3(std::min(1024, std::max(1, ceil_div(60L * n, 256)))), dim3(256), 0,
                       L->stream, (const BatchItem*)L->batch_items.p, n, L->batch_states.p, L->batch_reports.p);
    HIPCHK_CTX(L, hipGetLastError());
    std::vector<CgState> states(n);
    std::vector<double> reports(27 * (size_t)n);
    HIPCHK_CTX(L, hipMemcpyAsync(states.data(), L->batch_states.p, sizeof(CgState) * n, hipMemcpyDeviceToHost,
                                 L->stream));
    HIPCHK_CTX(L, d2h(L, reports.data(), L->batch_reports.p, sizeof(double) * 27 * n));
    for (int i = 0; i < n; i++) {
        mdn_ctx* c = ctxs[i];
        const CgState& S = states[i];
        const int nv = 3 * c->n_beads;
        if (coords[i])
            HIPCHK_CTX(L, hipMemcpyAsync(coords[i], c->x[S.parity].p, sizeof(double) * nv, hipMemcpyDeviceToHost,
                                         L->stream));
        if (forces_out && forces_out[i])
            HIPCHK_CTX(L, hipMemcpyAsync(forces_out[i], (S.fpar ? c->f2 : c->f).p, sizeof(double) * nv,
                                         hipMemcpyDeviceToHost, L->stream));
    }
    HIPCHK_CTX(L, hipStreamSynchronize(L->stream));
    int first_err = MDN_OK;
    for (int i = 0; i < n; i++) {
        mdn_ctx* c = ctxs[i];
        const CgState& S = states[i];
        const double* rep = reports.data() + 27L * i;
        c->last_cg = S;
        c->cur = S.parity;
        const bool ok = S.status == 0 && S.error_flags == 0;
        fill_result(S, rep, rep + 9, ok, &results[i]);
        int code = MDN_OK;
        if (S.status != 0) {
            code = (S.status == ST_NEED_FORCEN || S.status == 9) ? MDN_ERR_UNSUPPORTED : S.status;
            fail(c, code, S.status == ST_NEED_FORCEN ? "an excluded-volume pair needed the numerical force fallback, "
                                                       "which the batched minimizer does not run"
                          : S.status == 9 ? "more excluded-volume pairs need the Simpson energy than the capacity"
                                          : "minimization stopped with status %d", S.status);
        } else if (ok && rep[9 + 8] != 0.0) {
            code = MDN_ERR_UNSUPPORTED;
            fail(c, code, "excluded-volume pairs beyond the Simpson-energy capacity");
        } else if (S.error_flags & MDN_MIN_ERR_LINESEARCH) {
            code = MDN_ERR_LINESEARCH;
            fail(c, code, "line search failed (recovery disabled)");
        }
        if (code != MDN_OK && first_err == MDN_OK) {
            first_err = code;
            if (c != L) fail(L, code, "system %d: %s", i, c->err.c_str());
        }
    }
    return first_err;
}

// The schedule of the last batched minimization on a lead context: [workgroup size,
// largest-first order (1) or given order (0), tuned in that call (1) or reused (0)], and the
// tuning times (ms) of the candidates (128, 256, 512, 1024 threads, each in given then sorted order).
int mdn_last_batch_schedule(mdn_ctx* ctx, int32_t* out3, double* tune_ms) {
    if (!ctx || !out3) return MDN_ERR_ARG;
    out3[0] = ctx->batch_nt;
    out3[1] = ctx->batch_sorted;
    out3[2] = ctx->batch_tuned;
    if (tune_ms)
        for (int c = 0; c < MDN_BATCH_CANDIDATES; c++) tune_ms[c] = ctx->batch_tune_ms[c];
    return MDN_OK;
}

// Launch statistics of the last minimization: [line-search launches that did work, steps resumed
// after the host-run forceN fallback, force gathers, CG iterations].
int mdn_last_min_stats(mdn_ctx* ctx, int64_t* out) {
    if (!ctx || !out) return MDN_ERR_ARG;
    out[0] = ctx->last_cg.n_ls_launch;
    out[1] = ctx->last_cg.reserved[1];   // steps resumed after the host-run forceN fallback
    out[2] = ctx->last_cg.n_gather_launch;
    out[3] = ctx->last_cg.num_iter;
    return MDN_OK;
}

// ----------------------------------------------------------------------------------------
// Benchmark helpers: device-resident coordinates, HIP events on the context stream
// ----------------------------------------------------------------------------------------
extern "C++" {
template <typename R>
static int bench_forces_t(mdn_ctx* ctx, int32_t iters, double* ms_per_eval, double* kernel_ms) {
    const DevSys& s = ctx->sys;
    const double* x = ctx->x[ctx->cur].p;
    R* slots = reinterpret_cast<R*>(ctx->slots.p);
    const bool pieces = !ctx->bg_eval;
    const int nb = std::max(1, force_blocks(s, pieces));
    int st;
    // warm-up (and fallback check) once
    if ((st = eval_slots_t<R>(ctx, x, pieces, 3, nullptr))) return st;
    hipEvent_t e0, e1, k0, k1;
    HIPCHK(hipEventCreate(&e0)); HIPCHK(hipEventCreate(&e1)); HIPCHK(hipEventCreate(&k0)); HIPCHK(hipEventCreate(&k1));
    // whole evaluation: terms + gather
    HIPCHK(hipEventRecord(e0, ctx->stream));
    for (int i = 0; i < iters; i++) {
        hipLaunchKernelGGL(k_forces_terms<R>, dim3(nb), dim3(BLOCK), 0, ctx->stream, s, x, slots,
                           ctx->bad_count.p, ctx->bad_list.p, bead_block0(s, pieces), 3, (const double*)nullptr);
        launch_gather_plain<R>(ctx, s, 3, (const R*)slots, 0L, s.slot0[NFF], (const double*)nullptr, ctx->f.p, !pieces);
    }
    HIPCHK(hipEventRecord(e1, ctx->stream));
    // dominant kernel alone
    HIPCHK(hipEventRecord(k0, ctx->stream));
    for (int i = 0; i < iters; i++)
        hipLaunchKernelGGL(k_forces_terms<R>, dim3(nb), dim3(BLOCK), 0, ctx->stream, s, x, slots,
                           ctx->bad_count.p, ctx->bad_list.p, bead_block0(s, pieces), 3, (const double*)nullptr);
    HIPCHK(hipEventRecord(k1, ctx->stream));
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventSynchronize(k1));
    float a = 0, b = 0;
    HIPCHK(hipEventElapsedTime(&a, e0, e1));
    HIPCHK(hipEventElapsedTime(&b, k0, k1));
    if (ms_per_eval) *ms_per_eval = a / iters;
    if (kernel_ms) *kernel_ms = b / iters;
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipEventDestroy(k0); (void)hipEventDestroy(k1);
    return MDN_OK;
}
}  // extern "C++"

int mdn_bench_forces(mdn_ctx* ctx, int32_t precision, int32_t iters, double* ms_per_eval, double* kernel_ms) {
    int st = need_system(ctx, true);
    if (st) return st;
    if (precision != MDN_PREC_FP64 && precision != MDN_PREC_MIXED)
        return fail(ctx, MDN_ERR_ARG, "unknown precision %d", precision);
    if (iters <= 0) return fail(ctx, MDN_ERR_ARG, "iters must be positive");
    HIPCHK(hipSetDevice(ctx->device));
    return precision == MDN_PREC_MIXED ? bench_forces_t<float>(ctx, iters, ms_per_eval, kernel_ms)
                                       : bench_forces_t<double>(ctx, iters, ms_per_eval, kernel_ms);
}

// n_vec products (tangents at hb_v) as mdn_hvp_batch launches them, into hb_out.
static int hvp_batch_launch(mdn_ctx* ctx, const double* x, int prec, int n_vec, long slot_stride, long piece_stride) {
    HIPCHK(hipMemsetAsync(ctx->bad_count.p, 0, sizeof(int), ctx->stream));
    launch_tangents(ctx, x, prec, ctx->hb_v.p, n_vec, ctx->hb_slots.p, slot_stride, ctx->hb_piece.p, piece_stride);
    HIPCHK(hipGetLastError());
    gather_tangents(ctx, n_vec, ctx->hb_slots.p, slot_stride, ctx->hb_piece.p, piece_stride, ctx->hb_out.p);
    HIPCHK(hipGetLastError());
    return MDN_OK;
}

int mdn_bench_adjoint(mdn_ctx* ctx, int32_t precision, int32_t iters, const double* v, int32_t n_var, int32_t n_vec,
                      double* ms_per_hvp, double* ms_per_param_grad) {
    int st = need_system(ctx, true);
    if (st) return st;
    if (precision != MDN_PREC_FP64 && precision != MDN_PREC_MIXED)
        return fail(ctx, MDN_ERR_ARG, "unknown precision %d", precision);
    if (iters <= 0) return fail(ctx, MDN_ERR_ARG, "iters must be positive");
    if (!v || n_var != 3 * ctx->n_beads) return fail(ctx, MDN_ERR_ARG, "n_var must be 3*n_beads = %d", 3 * ctx->n_beads);
    if (n_vec < 1 || n_vec > MDN_HVP_BATCH_CHUNK)
        return fail(ctx, MDN_ERR_ARG, "n_vec must be in [1, %d]", MDN_HVP_BATCH_CHUNK);
    HIPCHK(hipSetDevice(ctx->device));
    const DevSys& s = ctx->sys;
    const double* x = ctx->x[ctx->cur].p;
    const long slot_stride = 3 * std::max(1L, s.slot0[NFF]) + GCHUNK_PAD;
    const long piece_stride = (long)NBOND * 3 * ctx->n_beads;
    if (n_vec == 1) {
        HIPCHK(ctx->tan_v.ensure(n_var));
        HIPCHK(hipMemcpyAsync(ctx->tan_v.p, v, sizeof(double) * n_var, hipMemcpyHostToDevice, ctx->stream));
    } else {
        HIPCHK(ctx->hb_v.ensure((size_t)n_vec * n_var));
        HIPCHK(ctx->hb_slots.ensure((size_t)n_vec * slot_stride));
        HIPCHK(ctx->hb_piece.ensure((size_t)n_vec * piece_stride));
        HIPCHK(ctx->hb_out.ensure((size_t)n_vec * n_var));
        HIPCHK(hipMemcpyAsync(ctx->hb_v.p, v, sizeof(double) * n_vec * n_var, hipMemcpyHostToDevice, ctx->stream));
    }
    auto hvps = [&]() {
        return n_vec == 1 ? hvp_launch(ctx, x, precision)
                          : hvp_batch_launch(ctx, x, precision, n_vec, slot_stride, piece_stride);
    };
    HIPCHK(ctx->ev_U.ensure(std::max(1, s.n[FF_EXVOL])));
    HIPCHK(ctx->grad_out.ensure(3 * (size_t)std::max(1, ctx->n_cyl)));
    // once with the checks: every pair has a tangent and a closed-form energy at x, so the timed
    // launches below do exactly the work of mdn_hvp (mdn_hvp_batch) and mdn_param_grad on the device
    if ((st = hvps())) return st;
    HIPCHK(hipMemcpyAsync(ctx->h_int, ctx->bad_count.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->h_int[0] != 0)
        return fail(ctx, MDN_ERR_UNSUPPORTED, "%d excluded-volume pairs need the numerical force fallback (no tangent)",
                    ctx->h_int[0]);
    if ((st = param_grad_launch(ctx, x, precision, true))) return st;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (s.n[FF_EXVOL] > 0 && ctx->h_int[0] != 0)
        return fail(ctx, MDN_ERR_UNSUPPORTED, "%d excluded-volume pairs need the Simpson energy", ctx->h_int[0]);
    hipEvent_t e0, e1, e2;
    HIPCHK(hipEventCreate(&e0)); HIPCHK(hipEventCreate(&e1)); HIPCHK(hipEventCreate(&e2));
    HIPCHK(hipEventRecord(e0, ctx->stream));
    for (int i = 0; i < iters && !st; i++) st = hvps();
    HIPCHK(hipEventRecord(e1, ctx->stream));
    for (int i = 0; i < iters && !st; i++) st = param_grad_launch(ctx, x, precision, false);
    HIPCHK(hipEventRecord(e2, ctx->stream));
    HIPCHK(hipEventSynchronize(e2));
    float a = 0, b = 0;
    HIPCHK(hipEventElapsedTime(&a, e0, e1));
    HIPCHK(hipEventElapsedTime(&b, e1, e2));
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipEventDestroy(e2);
    if (st) return st;
    if (ms_per_hvp) *ms_per_hvp = a / iters / n_vec;
    if (ms_per_param_grad) *ms_per_param_grad = b / iters;
    return MDN_OK;
}

int mdn_bench_energy(mdn_ctx* ctx, int32_t precision, int32_t iters, double* ms_per_eval) {
    int st = need_system(ctx, true);
    if (st) return st;
    if (precision != MDN_PREC_FP64 && precision != MDN_PREC_MIXED)
        return fail(ctx, MDN_ERR_ARG, "unknown precision %d", precision);
    if (iters <= 0) return fail(ctx, MDN_ERR_ARG, "iters must be positive");
    HIPCHK(hipSetDevice(ctx->device));
    const double* x = ctx->x[ctx->cur].p;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0)); HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, ctx->stream));
    for (int i = 0; i < iters; i++) launch_energy(ctx, x, precision);
    HIPCHK(hipEventRecord(e1, ctx->stream));
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventSynchronize(e1));
    float a = 0;
    HIPCHK(hipEventElapsedTime(&a, e0, e1));
    if (ms_per_eval) *ms_per_eval = a / iters;
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
    return MDN_OK;
}

}  // extern "C"
