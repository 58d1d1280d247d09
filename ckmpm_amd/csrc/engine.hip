// Host orchestration of the CK-MPM substep and the C-ABI of libckmpm.so.
//
// One engine = one HIP stream.  Every per-substep quantity that the reference reads
// back to the host (block counts, bin offsets, max |v|) stays in device memory:
// kernels read the counts they need, so a substep is enqueued without a host sync
// (the reference needs about six, mpm_engine.cuh:537-542,650,724-737,783,803-807,846-850).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "engine_impl.h"
#include "setup_wide.h"

using namespace ckmpm;

namespace {



uint32_t GridFor(uint32_t blocks) {
    uint32_t g = std::max<uint32_t>(blocks, 8u);
    g = std::min<uint32_t>(g, 16384u);
    return (g + 7u) & ~7u;
}

int FlagsToCode(uint32_t f) {
    if (f & kFlagNonFinite) return CKMPM_ERR_NONFINITE;
    if (f & kFlagCapacity) return CKMPM_ERR_CAPACITY;
    if (f & kFlagOutOfDomain) return CKMPM_ERR_OUT_OF_DOMAIN;
    if (f & (kFlagOutOfWindow | kFlagMissingBlock)) return CKMPM_ERR_OUT_OF_WINDOW;
    return CKMPM_OK;
}

int ReadStatus(ckmpm_engine* e, Status& st, LayoutCounts& lc) {
    CK_HIP(e, hipMemcpyAsync(&st, e->status, sizeof(Status), hipMemcpyDeviceToHost, e->stream));
    CK_HIP(e, hipMemcpyAsync(&lc, e->L[e->cur].lc, sizeof(LayoutCounts), hipMemcpyDeviceToHost, e->stream));
    CK_HIP(e, hipStreamSynchronize(e->stream));
    e->a_hint = lc.A;
    return CKMPM_OK;
}

// Launches of a substep recorded for a fused batch instead of enqueued: per (stage,
// kernel, variant), the argument structs and workgroup counts of the scenes in the order
// they were recorded.  Stages follow the substep's dependency order (grid update, G2P2G,
// the four rebin kernels, the two extension passes' mark and emit); launches of one stage
// with different kernels or variants are independent of each other.
struct FusedPhase {
    int stage = 0, kernel = 0, variant = 0;
    std::vector<uint8_t> args;    // FusedArgsSize(kernel) bytes per scene
    std::vector<uint32_t> grids;  // workgroups per scene
};
struct FusedRec {
    std::vector<FusedPhase> ph;
    bool shared = false;  // several engines: their G2P2G launches run side by side
    void Add(int stage, int kernel, int variant, const void* a, size_t size, uint32_t grid) {
        FusedPhase* p = nullptr;
        for (FusedPhase& q : ph)
            if (q.stage == stage && q.kernel == kernel && q.variant == variant) p = &q;
        if (!p) {
            ph.emplace_back();
            p = &ph.back();
            p->stage = stage;
            p->kernel = kernel;
            p->variant = variant;
        }
        const uint8_t* b = static_cast<const uint8_t*>(a);
        p->args.insert(p->args.end(), b, b + size);
        p->grids.push_back(grid);
    }
};
enum Stage : int { kStGrid, kStG2P2G, kStCount, kStScan, kStScatter, kStMark, kStEmit };

// Launch now on s, or record for a fused batch.
template <class T>
void Run(FusedRec* rec, int stage, int kernel, int variant, const T& a, uint32_t grid, hipStream_t s) {
    if (rec) rec->Add(stage, kernel, variant, &a, sizeof(T), grid);
    else LaunchOne(kernel, variant, &a, grid, s);
}

// Neighbour and exterior extension of layout L (its A keys already written, and in
// the table unless insert_a), on stream s.  The map tile counts must be zero: a substep's
// rebin count kernel clears them, the setup before its call.
void ExtendLayout(ckmpm_engine* e, LayoutBuf& L, bool insert_a, hipStream_t s, FusedRec* rec = nullptr) {
    if (insert_a) LaunchTableInsertA(L.keys, L.lc, e->cap, L.table, s);
    // marking grids sized for the expected block count (the capacity before the first read)
    const uint32_t hint = e->a_hint ? e->a_hint + e->a_hint / 4 + 64u : e->cap;
    const MarkArgs m{e->dom, L.keys, L.lc, e->cap, L.table, e->ext_map, e->ext_cnt, e->ext_tile_shift};
    Run(rec, kStMark, kFkMark, 0, m, MarkGrid(e->cap, hint), s);
    const EmitArgs em{e->ext_map, e->ext_cnt, e->ext_tiles, L.keys, L.table, L.lc, e->cap, e->status};
    Run(rec, kStEmit, kFkEmit, e->emit_variant, em, e->ext_tiles, s);
}

void Mark(ckmpm_engine* e) {
    if (!e->profile) return;
    if (e->ev_used == e->events.size()) {
        hipEvent_t ev;
        if (hipEventCreate(&ev) != hipSuccess) return;
        e->events.push_back(ev);
    }
    (void)hipEventRecord(e->events[e->ev_used++], e->stream);
}

// Grid update with dt: velocities of the current layout from the previous layout's
// mass / momentum; clears the next P2G target.
// The grid's max speed (the CFL dt, the fixed-point range) leaves out nodes holding less
// than 2^-20 of the lightest particle's mass (see GridUpdateArgs::vmax_mass_floor).
float VmaxMassFloor(const ckmpm_engine* e) {
    float mmin = 0.0f;
    for (const ModelBuf& m : e->models) mmin = (mmin == 0.0f || m.mat.mass < mmin) ? m.mat.mass : mmin;
    return std::ldexp(mmin, -20);
}

// A substep swaps the roles of the double-buffered layouts, particle states and grids.
void FlipRoles(ckmpm_engine* e) {
    e->cur = 1 - e->cur;
    e->sc = 1 - e->sc;
    e->mprev = 1 - e->mprev;
}

void PhaseGrid(ckmpm_engine* e, double dt, FusedRec* rec = nullptr) {
    hipStream_t s = e->stream;
    Mark(e);
    GridUpdateArgs g{};
    g.dom = e->dom;
    g.keys = e->L[e->cur].keys;
    g.lc = e->L[e->cur].lc;
    g.table_prev = e->L[1 - e->cur].table;
    g.lc_prev = e->L[1 - e->cur].lc;
    g.M_prev = e->M[e->mprev];
    g.V = e->V;
    g.M_zero = e->M[1 - e->mprev];
    g.dt = static_cast<float>(dt);
    g.gravity = e->sim.gravity;
    g.mass_clamp = e->sim.mass_clamp;
    g.vmax_mass_floor = VmaxMassFloor(e);
    g.cap = e->cap;
    g.st = e->status;
    g.nmodels = static_cast<int>(e->models.size());
    for (int i = 0; i < g.nmodels; ++i) g.blkmax[i] = e->models[i].blkmax;
    g.bc = e->bc;
    g.custom = e->bc_custom ? 1 : 0;
    g.slot = static_cast<int>(e->substeps & 1u);
    // 8 blocks x 4 chains per workgroup and pass: 0.047 A workgroups cover Nb (about 1.2 A)
    Run(rec, kStGrid, kFkGrid, g.custom, g, GridFor(e->a_hint * 3u / 64u + 8u), s);
    Mark(e);
}

// G2P2G(dt, next_dt) for every model, then the rebin into the next layout.
void PhaseParticles(ckmpm_engine* e, double dt, double next_dt, FusedRec* rec = nullptr) {
    hipStream_t s = e->stream;
    LayoutBuf& C = e->L[e->cur];
    LayoutBuf& P = e->L[1 - e->cur];
    float* Mout = e->M[1 - e->mprev];
    const int sc = e->sc, sn = 1 - e->sc;
    // G2P2G workgroups, each taking particle blocks round-robin: 1.25 per block for one
    // model alone; half a workgroup per block where launches run side by side in graph
    // replays (several models, or the fused launches of a batch of engines), so that the
    // launches share the CUs.  On the 8M fluid + sand scene 0.5 per block took the
    // substep from 0.654 to 0.628 ms (0.625: 0.637; 0.31: 0.629 but slower as plain
    // launches; 0.25: 0.635), the two-model 30k spheres gained 5% and the 64-scene batch
    // 1-2% (0.25: no more); the single-model 1M jelly lost 12% (one engine in a graph).
    const bool side_by_side = e->models.size() > 1 || (rec && rec->shared);
    const uint32_t grid = GridFor((side_by_side ? e->a_hint / 2u : e->a_hint + e->a_hint / 4u) + 8u);
    // The models' launches run one after the other on the engine stream.  Running the
    // second model's on a second stream, concurrently with the first (the launches are
    // independent; their grid deposits are commutative atomic adds), measured slower on
    // the 8M fluid + sand scene: 0.614 against 0.607 ms.
    for (size_t mi = 0; mi < e->models.size(); ++mi) {
        ModelBuf& m = e->models[mi];
        G2P2GArgs a{};
        a.dom = e->dom;
        a.mat = m.mat;
        a.keys = C.keys;
        a.table = C.table;
        a.lc = C.lc;
        a.V = e->V;
        a.M = Mout;
        a.st_in = m.st[sc];
        a.st_out = m.st[sn];
        a.n = m.n;
        a.perm = m.perm[sc];
        a.off = m.off[sc];
        a.code = m.code;
        a.cnt = m.cnt;
        a.pp_attr = m.pp_attr;
        a.rigid_attr = m.rigid_attr;
        for (int k = 0; k < 3; ++k) a.rigid_v[k] = m.rigid_v[k];
        a.dt = static_cast<float>(dt);
        a.next_dt = static_cast<float>(next_dt);
        a.st = e->status;
        a.model = static_cast<int>(mi);
        a.slot = static_cast<int>(e->substeps & 1u);
        a.blkmax = m.blkmax;
        AdjointIdBuffers(e, mi, &a.ids_in, &a.ids_out);
        const bool tune = e->tuning && !rec;
        if (tune) (void)hipEventRecord(m.tune_ev[0], s);
        Run(rec, kStG2P2G, kFkG2P2G, m.mat.kind * 2 + (m.threads == 512 ? 1 : 0), a, G2P2GGrid(grid), s);
        if (tune) (void)hipEventRecord(m.tune_ev[1], s);
    }
    Mark(e);
    // ---- rebin: new layout in P (the previous layout's buffers are free now)
    RebinModels rm{};
    rm.nmodels = static_cast<int>(e->models.size());
    for (int i = 0; i < rm.nmodels; ++i) {
        ModelBuf& m = e->models[i];
        rm.cnt[i] = m.cnt;
        rm.cnt_e[i] = m.cnt_e;
        rm.eoff[i] = m.eoff;
        rm.base[i] = m.base;
        rm.off_next[i] = m.off[sn];
        rm.off_cur[i] = m.off[sc];
        rm.code[i] = m.code;
        rm.perm_next[i] = m.perm[sn];
        rm.n[i] = m.n;
    }
    // exterior blocks are about 1.4 A; the kernels loop over any count
    const RebinCountArgs ca{e->dom, C.keys, C.table, C.lc, rm, e->tsum, P.keys, P.lc, P.table, e->cap,
                            e->ext_cnt, 2u * e->ext_tiles};
    Run(rec, kStCount, kFkRebinCount, 0, ca, RebinCountGrid(e->cap, 2u * e->a_hint + 64u), s);
    const RebinScanArgs sa{C.keys, C.lc, rm, e->tsum, P.keys, P.table, P.lc, e->cap, e->status};
    Run(rec, kStScan, kFkRebinScan, 0, sa, RebinScanGrid(e->cap), s);
    // Two independent chains from here: the particle permutation (RebinScatter: current
    // layout, orderings) and the extension of the next layout (its neighbour and exterior
    // blocks, map, tile counts).  The scatter and the extension's mark pass share one
    // launch (the first workgroups mark), then the emit pass runs.
    const uint32_t hint = e->a_hint ? e->a_hint + e->a_hint / 4 + 64u : e->cap;
    ScatterMarkArgs sm{};
    sm.x = RebinScatterArgs{e->dom, C.keys, C.table, C.lc, rm, e->cap};
    sm.m = MarkArgs{e->dom, P.keys, P.lc, e->cap, P.table, e->ext_map, e->ext_cnt, e->ext_tile_shift};
    sm.mark_wgs = MarkGrid(e->cap, hint);
    // a wavefront per particle block, 4 per workgroup
    Run(rec, kStScatter, kFkScatterMark, 0, sm, sm.mark_wgs + GridFor((e->a_hint + e->a_hint / 4) / 4u + 8u), s);
    const EmitArgs em{e->ext_map, e->ext_cnt, e->ext_tiles, P.keys, P.table, P.lc, e->cap, e->status};
    Run(rec, kStEmit, kFkEmit, e->emit_variant, em, e->ext_tiles, s);
    Mark(e);
    FlipRoles(e);
    e->substeps++;
}

// A copy ordered after the work queued on the engine's stream, waited for.
int CopyOnStream(ckmpm_engine* e, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    CK_HIP(e, hipMemcpyAsync(dst, src, bytes, kind, e->stream));
    CK_HIP(e, hipStreamSynchronize(e->stream));
    return CKMPM_OK;
}

int CheckReady(ckmpm_engine* e) {
    if (!e) return CKMPM_ERR_INVALID;
    if (!e->setup_done) return Fail(e, CKMPM_ERR_INVALID, "call ckmpm_initial_setup first");
    return CKMPM_OK;
}

}  // namespace

extern "C" {

int ckmpm_create(const ckmpm_domain_desc* domain, const ckmpm_sim_desc* sim, ckmpm_engine** out) {
    if (!domain || !sim || !out) return CKMPM_ERR_INVALID;
    *out = nullptr;
    // descriptors first (no device needed to reject them)
    for (int a = 0; a < 3; ++a)
        if (domain->dim[a] <= 0 || domain->dim[a] > 1024) return CKMPM_ERR_INVALID;
    if (!(sim->dx > 0.0f) || sim->block_capacity == 0) return CKMPM_ERR_INVALID;
    // at most 1024^3 = 2^30 blocks: a block key fits 32 bits with room for the empty marker
    const uint64_t nblocks = static_cast<uint64_t>(domain->dim[0]) * domain->dim[1] * domain->dim[2];
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CKMPM_ERR_NO_DEVICE;
    ckmpm_engine* e = new ckmpm_engine();
    for (int a = 0; a < 3; ++a) {
        e->dom.dim[a] = domain->dim[a];
        e->dom.off[a] = domain->offset[a];
    }
    e->dom.nblocks = static_cast<uint32_t>(nblocks);
    e->dom.dx = sim->dx;
    e->dom.inv_dx = static_cast<float>(1.0 / static_cast<double>(sim->dx));  // mpm_config.h:22-25
    e->sim = *sim;
    e->cap = std::min<uint32_t>(sim->block_capacity, e->dom.nblocks);
    e->dt_default = sim->dt;
    if (sim->fps) e->dt_default = std::min<double>(e->dt_default, 1.0 / sim->fps);  // mpm_engine.cuh:178
    int rc = CKMPM_OK;
    auto check = [&](int r) { if (rc == CKMPM_OK) rc = r; };
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) rc = Fail(e, CKMPM_ERR_HIP, "stream");
    for (int i = 0; i < 2 && rc == CKMPM_OK; ++i) {
        check(Alloc(e, &e->L[i].keys, e->cap));
        check(Alloc(e, &e->L[i].table, nblocks));
        check(Alloc(e, &e->L[i].lc, 1));
        check(Alloc(e, &e->M[i], static_cast<size_t>(e->cap) * kMFloats));
    }
    e->nwords = static_cast<uint32_t>((nblocks + 31) / 32);
    check(Alloc(e, &e->V, static_cast<size_t>(e->cap) * kVFloats));
    check(Alloc(e, &e->status, 1));
    check(Alloc(e, &e->bc, 1));
    check(Alloc(e, &e->bitmap, e->nwords));
    e->ext_tiles = ExtTiles(nblocks);
    e->ext_tile_words = ExtTileWords(nblocks);
    e->ext_tile_shift = ExtTileShift(nblocks);
    e->emit_variant = EmitVariant(nblocks);
    check(Alloc(e, &e->ext_map, static_cast<size_t>(e->ext_tiles) * e->ext_tile_words));
    check(Alloc(e, &e->ext_cnt, 2 * static_cast<size_t>(e->ext_tiles)));
    check(Alloc(e, &e->tsum, RebinTileSumWords(e->cap)));
    check(Alloc(e, &e->word_off, e->nwords));
    check(Alloc(e, &e->total, 4));
    check(Alloc(e, &e->scan_tmp, ScanTempWords(std::max<size_t>(e->cap, e->nwords))));
    if (rc == CKMPM_OK) {
        for (int i = 0; i < 2; ++i) {
            (void)hipMemsetAsync(e->L[i].table, 0xFF, nblocks * sizeof(uint32_t), e->stream);
            (void)hipMemsetAsync(e->L[i].lc, 0, sizeof(LayoutCounts), e->stream);
        }
        (void)hipMemsetAsync(e->bitmap, 0, static_cast<size_t>(e->nwords) * sizeof(uint32_t), e->stream);
        (void)hipMemsetAsync(e->ext_map, 0, static_cast<size_t>(e->ext_tiles) * e->ext_tile_words * sizeof(uint32_t),
                             e->stream);
        (void)hipMemsetAsync(e->status, 0, sizeof(Status), e->stream);
        (void)hipMemsetAsync(e->bc, 0, sizeof(BoundaryDev), e->stream);
        if (hipStreamSynchronize(e->stream) != hipSuccess) rc = Fail(e, CKMPM_ERR_HIP, "init");
    }
    if (rc != CKMPM_OK) {
        ckmpm_destroy(e);
        return rc;
    }
    *out = e;
    return CKMPM_OK;
}

int ckmpm_add_model(ckmpm_engine* e, const ckmpm_material_desc* mat, float mass, float volume, const float* xyz,
                    uint64_t n, const float* vel, uint64_t nvel, const float* lambda_p, const float* mu_p,
                    int32_t* model_index) {
    if (!e || !mat || !xyz || !vel) return CKMPM_ERR_INVALID;
    if (e->setup_done) return Fail(e, CKMPM_ERR_INVALID, "models must be added before ckmpm_initial_setup");
    if (e->models.size() >= static_cast<size_t>(kMaxModels)) return Fail(e, CKMPM_ERR_INVALID, "too many models");
    if (n == 0 || n >= (1ull << 31)) return Fail(e, CKMPM_ERR_INVALID, "particle count");
    if (nvel != 1 && nvel != n) return Fail(e, CKMPM_ERR_INVALID, "velocity count must be 1 or n");
    if (mat->kind < 0 || mat->kind > 5) return Fail(e, CKMPM_ERR_INVALID, "material kind");
    if ((lambda_p == nullptr) != (mu_p == nullptr)) return Fail(e, CKMPM_ERR_INVALID, "lambda_p and mu_p go together");
    if (lambda_p && mat->kind == kFluid) return Fail(e, CKMPM_ERR_INVALID, "per-particle Lame parameters need a solid");
    ModelBuf m;
    m.kind = mat->kind;
    m.nattr = AttrCount(m.kind);
    if (lambda_p) {
        m.pp_attr = m.nattr;  // two extra state attributes travel with the particle
        m.nattr += 2;
        m.lam0.assign(lambda_p, lambda_p + n);
        m.mu0.assign(mu_p, mu_p + n);
    }
    m.n = static_cast<uint32_t>(n);
    m.mat.kind = mat->kind;
    m.mat.lambda = mat->lambda;
    m.mat.mu = mat->mu;
    m.mat.bulk = mat->bulk;
    m.mat.gamma = mat->gamma;
    m.mat.gamma_int = (mat->gamma >= 1.0f && mat->gamma <= 16.0f && mat->gamma == std::floor(mat->gamma))
                          ? static_cast<int>(mat->gamma)
                          : 0;
    m.mat.viscosity = mat->viscosity;
    m.mat.J0 = mat->J0;
    m.mat.sin_phi = static_cast<float>(std::sin(static_cast<double>(mat->friction_angle_deg) * 3.14159265358979323846 / 180.0));
    m.mat.yield = mat->yield_stress;
    m.mat.beta = mat->beta;
    m.mat.xi = mat->hardening_factor;
    m.mat.msq = mat->m * mat->m;
    m.mat.alpha0 = mat->alpha0;
    m.mat.hardening = mat->hardening_on != 0;
    m.mat.mass = mass;
    m.mat.volume = volume;
    m.x0.assign(xyz, xyz + 3 * n);
    m.v0.assign(vel, vel + 3 * nvel);
    m.nvel = static_cast<uint32_t>(nvel);
    // default dt from the initial speeds: mpm_engine.cuh:109-117
    double vmax = 0.0;
    for (uint64_t i = 0; i < nvel; ++i) {
        const double s = std::sqrt(double(vel[3 * i]) * vel[3 * i] + double(vel[3 * i + 1]) * vel[3 * i + 1] +
                                   double(vel[3 * i + 2]) * vel[3 * i + 2]);
        vmax = std::max(vmax, s);
    }
    if (vmax > 0.0 && e->sim.cfl > 0.0f) e->dt_default = std::min(e->dt_default, double(e->sim.dx) * e->sim.cfl / vmax);
    const size_t N = m.n;
    for (int i = 0; i < 2; ++i) {
        CK_TRY(Alloc(e, &m.st[i], N * m.nattr));
        CK_TRY(Alloc(e, &m.perm[i], N));
        CK_TRY(Alloc(e, &m.off[i], static_cast<size_t>(e->cap) + 1));
    }
    CK_TRY(Alloc(e, &m.code, N));
    CK_TRY(Alloc(e, &m.cnt, static_cast<size_t>(e->cap) * kDirections));
    CK_TRY(Alloc(e, &m.base, static_cast<size_t>(e->cap) * kDirections));
    CK_TRY(Alloc(e, &m.cnt_e, e->cap));
    CK_TRY(Alloc(e, &m.eoff, e->cap));
    CK_TRY(Alloc(e, &m.blkmax, e->cap));
    CK_HIP(e, hipMemsetAsync(m.blkmax, 0, static_cast<size_t>(e->cap) * sizeof(float), e->stream));
    e->models.push_back(std::move(m));
    if (model_index) *model_index = static_cast<int32_t>(e->models.size() - 1);
    return CKMPM_OK;
}

int ckmpm_set_rigid(ckmpm_engine* e, int32_t model, uint64_t count, const float velocity[3]) {
    if (!e || !velocity) return CKMPM_ERR_INVALID;
    if (model < 0 || model >= static_cast<int32_t>(e->models.size())) return Fail(e, CKMPM_ERR_INVALID, "model index");
    ModelBuf& m = e->models[model];
    if (m.kind == kFluid && count) return Fail(e, CKMPM_ERR_INVALID, "rigid particles need a solid model");
    if (count > m.n) return Fail(e, CKMPM_ERR_INVALID, "more rigid particles than the model has");
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(velocity[k])) return Fail(e, CKMPM_ERR_INVALID, "rigid velocity must be finite");
    if (e->setup_done) {
        if (count != m.rigid_count) return Fail(e, CKMPM_ERR_INVALID, "the rigid count is fixed at ckmpm_initial_setup");
    } else if (count && !m.rigid_attr) {
        if (e->adj) return Fail(e, CKMPM_ERR_INVALID, "the adjoint does not differentiate rigid particles");
        // one more state attribute: replace the two state buffers
        for (int i = 0; i < 2; ++i) {
            auto it = std::find(e->allocs.begin(), e->allocs.end(), static_cast<void*>(m.st[i]));
            if (it != e->allocs.end()) e->allocs.erase(it);
            CK_HIP(e, hipFree(m.st[i]));
            m.st[i] = nullptr;
        }
        m.rigid_attr = m.nattr;
        m.nattr += 1;
        for (int i = 0; i < 2; ++i) CK_TRY(Alloc(e, &m.st[i], static_cast<size_t>(m.n) * m.nattr));
    }
    m.rigid_count = count;
    bool changed = false;
    for (int k = 0; k < 3; ++k) {
        changed |= m.rigid_v[k] != velocity[k];
        m.rigid_v[k] = velocity[k];
    }
    if (changed) ++e->args_version;
    return CKMPM_OK;
}

int ckmpm_initial_setup(ckmpm_engine* e) {
    if (!e) return CKMPM_ERR_INVALID;
    if (e->setup_done) return Fail(e, CKMPM_ERR_INVALID, "setup already done");
    if (e->models.empty()) return Fail(e, CKMPM_ERR_INVALID, "no models");
    hipStream_t s = e->stream;
    uint32_t nmax = 0;
    for (ModelBuf& m : e->models) nmax = std::max(nmax, m.n);
    float *xyz = nullptr, *vel = nullptr, *lam = nullptr, *mu = nullptr;
    uint32_t* keys = nullptr;
    std::vector<uint32_t*> sorted(e->models.size(), nullptr);
    const size_t mark = e->allocs.size();
    CK_TRY(Alloc(e, &xyz, 3 * static_cast<size_t>(nmax)));
    CK_TRY(Alloc(e, &vel, 3 * static_cast<size_t>(nmax)));
    CK_TRY(Alloc(e, &lam, nmax));
    CK_TRY(Alloc(e, &mu, nmax));
    SetupTemps tmp;
    for (uint32_t*& t : tmp.t) CK_TRY(Alloc(e, &t, nmax));
    CK_TRY(Alloc(e, &tmp.keys_alt, nmax));
    CK_TRY(Alloc(e, &tmp.vals_alt, nmax));
    CK_TRY(Alloc(e, &tmp.radix, std::max(RadixTempWords(nmax), ScanTempWords(nmax))));
    int key_bits = 1;
    while ((1ull << key_bits) < e->dom.nblocks) ++key_bits;
    LayoutBuf& L = e->L[0];
    // particle blocks from FP32 positions; per-model canonical setup ordering
    for (size_t i = 0; i < e->models.size(); ++i) {
        ModelBuf& m = e->models[i];
        CK_TRY(Alloc(e, &keys, m.n));
        sorted[i] = keys;
        uint32_t* vals = m.perm[0];
        CK_HIP(e, hipMemcpyAsync(xyz, m.x0.data(), 3 * sizeof(float) * m.n, hipMemcpyHostToDevice, s));
        // the cell key block << 6 | cell fits 32 bits up to 2^26 blocks; larger domains
        // sort by cell and block in two passes
        if (key_bits + 6 <= 32) SetupOrder(e->dom, xyz, m.n, key_bits, keys, vals, tmp, e->bitmap, e->status, s);
        else SetupOrderWide(e->dom, xyz, m.n, key_bits, keys, vals, tmp, e->bitmap, e->status, s);
        AdjointOnSetupOrder(e, i, vals, xyz);
        if (m.pp_attr) {
            CK_HIP(e, hipMemcpyAsync(lam, m.lam0.data(), sizeof(float) * m.n, hipMemcpyHostToDevice, s));
            CK_HIP(e, hipMemcpyAsync(mu, m.mu0.data(), sizeof(float) * m.n, hipMemcpyHostToDevice, s));
        }
        LaunchInitState(e->dom, m.kind, xyz, lam, mu, m.pp_attr, vals, m.n, InitAttr(m.mat), m.st[0], s, m.rigid_attr,
                        m.rigid_count);
        CK_HIP(e, hipStreamSynchronize(s));
    }
    ExclusiveScan(e->bitmap, e->word_off, nullptr, e->nwords, e->nwords, e->total, true, e->scan_tmp, s);
    LaunchBitmapToKeysA(e->bitmap, e->nwords, e->word_off, e->total, L.keys, L.lc, e->cap, e->status, s);
    for (size_t i = 0; i < e->models.size(); ++i)
        LaunchSetupOffsets(sorted[i], e->models[i].n, L.keys, L.lc, e->cap, e->models[i].off[0], s);
    CK_HIP(e, hipMemsetAsync(e->ext_cnt, 0, 2 * static_cast<size_t>(e->ext_tiles) * sizeof(uint32_t), s));
    ExtendLayout(e, L, true, s);
    CK_HIP(e, hipMemsetAsync(e->M[0], 0, static_cast<size_t>(e->cap) * kMFloats * sizeof(float), s));
    for (ModelBuf& m : e->models) {
        CK_HIP(e, hipMemcpyAsync(xyz, m.x0.data(), 3 * sizeof(float) * m.n, hipMemcpyHostToDevice, s));
        CK_HIP(e, hipMemcpyAsync(vel, m.v0.data(), 3 * sizeof(float) * m.nvel, hipMemcpyHostToDevice, s));
        LaunchRasterize(e->dom, xyz, vel, m.n, m.nvel, m.mat.mass, L.table, L.lc, e->M[0], e->status, s);
        CK_HIP(e, hipStreamSynchronize(s));
    }
    // previous layout = setup layout (the first grid update reads M[0] through it)
    CK_HIP(e, hipMemcpyAsync(e->L[1].keys, L.keys, e->cap * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    CK_HIP(e, hipMemcpyAsync(e->L[1].table, L.table, static_cast<size_t>(e->dom.nblocks) * sizeof(uint32_t),
                             hipMemcpyDeviceToDevice, s));
    CK_HIP(e, hipMemcpyAsync(e->L[1].lc, L.lc, sizeof(LayoutCounts), hipMemcpyDeviceToDevice, s));
    // the stored state was written in the setup ordering: it is also the "previous" one
    for (ModelBuf& m : e->models)
        CK_HIP(e, hipMemcpyAsync(m.off[1], m.off[0], (static_cast<size_t>(e->cap) + 1) * sizeof(uint32_t),
                                 hipMemcpyDeviceToDevice, s));
    e->cur = 0;
    e->mprev = 0;
    e->sc = 0;
    Status st;
    LayoutCounts lc;
    CK_TRY(ReadStatus(e, st, lc));
    // release the setup temporaries
    for (size_t i = mark; i < e->allocs.size(); ++i) (void)hipFree(e->allocs[i]);
    e->allocs.resize(mark);
    for (ModelBuf& m : e->models) {
        std::vector<float>().swap(m.x0);
        std::vector<float>().swap(m.v0);
        std::vector<float>().swap(m.lam0);
        std::vector<float>().swap(m.mu0);
    }
    e->setup_done = true;
    e->time = 0.0;
    e->cur_dt = 0.0;
    const int code = FlagsToCode(st.flags);
    if (code != CKMPM_OK) return Fail(e, code, "initial setup flagged 0x" + std::to_string(st.flags));
    AdjointAfterSetup(e);
    return CKMPM_OK;
}

int ckmpm_substep(ckmpm_engine* e, double dt, double next_dt) {
    CK_TRY(CheckReady(e));
    CK_TRY(AdjointCheckTape(e));
    CK_TRY(ConservationCheck(e));
    PhaseGrid(e, dt);
    ConservationAfterGrid(e);
    // recorded after the grid phase, which leaves every recorded input unchanged and
    // provides the grid velocities of the substep
    CK_TRY(AdjointBeforeSubstep(e, dt, next_dt));
    PhaseParticles(e, dt, next_dt);
    AdjointAfterSubstep(e);
    // simulated time as the frame loop keeps it (mpm_engine.cuh:870-871: the timestamp
    // advances by the next substep's dt)
    e->cur_dt = next_dt;
    e->time += next_dt;
    const hipError_t r = hipGetLastError();
    if (r != hipSuccess) return Fail(e, CKMPM_ERR_HIP, std::string("launch: ") + hipGetErrorString(r));
    return CKMPM_OK;
}

int ckmpm_set_boundary(ckmpm_engine* e, const ckmpm_boundary_desc* b) {
    if (!e) return CKMPM_ERR_INVALID;
    BoundaryDev h{};
    if (b && e->adj)
        return Fail(e, CKMPM_ERR_INVALID, "the adjoint differentiates the default boundary only");
    if (b) {
        if (b->walls < CKMPM_WALLS_STICKY || b->walls > CKMPM_WALLS_NONE)
            return Fail(e, CKMPM_ERR_INVALID, "walls must be CKMPM_WALLS_STICKY, _SLIP or _NONE");
        if (b->nregions > CKMPM_MAX_REGIONS) return Fail(e, CKMPM_ERR_INVALID, "more than CKMPM_MAX_REGIONS regions");
        for (uint32_t r = 0; r < b->nregions; ++r)
            if (b->regions[r].kind < CKMPM_REGION_FIXED || b->regions[r].kind > CKMPM_REGION_ROTATION)
                return Fail(e, CKMPM_ERR_INVALID, "region " + std::to_string(r) + ": unknown kind");
        h.custom = 1;
        h.d = *b;
    }
    CK_HIP(e, hipMemcpyAsync(e->bc, &h, sizeof(h), hipMemcpyHostToDevice, e->stream));
    CK_HIP(e, hipStreamSynchronize(e->stream));
    e->bc_custom = b != nullptr;
    return CKMPM_OK;
}

// Simulate's frame loop (mpm_engine.cuh:514-876) with its dt rules.
int ckmpm_run_frames(ckmpm_engine* e, uint32_t frames) {
    CK_TRY(CheckReady(e));
    // a frame lasts 1 / fps: without a frame rate the loop below would never end
    if (e->sim.fps == 0) return Fail(e, CKMPM_ERR_INVALID, "ckmpm_run_frames needs sim.fps > 0");
    const double spf = 1.0 / e->sim.fps;
    for (uint32_t f = 0; f < frames; ++f) {
        const double frame_end = e->time + spf;
        while (e->time < frame_end) {
            CK_TRY(AdjointCheckTape(e));
            CK_TRY(ConservationCheck(e));
            PhaseGrid(e, e->cur_dt);
            ConservationAfterGrid(e);
            Status st;
            LayoutCounts lc;
            CK_TRY(ReadStatus(e, st, lc));
            const float vmax = std::sqrt(__builtin_bit_cast(float, st.vmax2_bits[e->substeps & 1u]));
            if (std::isinf(vmax) || (st.flags & kFlagNonFinite))
                return Fail(e, CKMPM_ERR_NONFINITE, "velocity reached infinity");
            double next_dt = std::min(e->dt_default, frame_end - e->time);
            if (vmax > 0.0f) next_dt = std::min(next_dt, double(e->sim.dx) * e->sim.cfl / vmax);
            CK_TRY(AdjointBeforeSubstep(e, e->cur_dt, next_dt));
            PhaseParticles(e, e->cur_dt, next_dt);
            AdjointAfterSubstep(e);
            e->cur_dt = next_dt;
            e->time += e->cur_dt;
        }
    }
    return ckmpm_synchronize(e);
}

// Autotuner: each candidate G2P2G workgroup size runs `substeps` real substeps (the
// simulation advances as with ckmpm_substep at the engine's dt); each model keeps the
// size with the smallest total G2P2G device time.
int ckmpm_autotune(ckmpm_engine* e, uint32_t substeps) {
    CK_TRY(CheckReady(e));
    if (substeps == 0) return CKMPM_OK;
    static const int kCandidates[2] = {256, 512};
    const size_t nm = e->models.size();
    for (ModelBuf& m : e->models)
        for (hipEvent_t& ev : m.tune_ev)
            if (!ev) CK_HIP(e, hipEventCreate(&ev));
    // Rounds of one substep per candidate, interleaved so that both see the same scene
    // state; round 0 only warms both kernels up (first launches, caches) and is not timed.
    std::vector<std::vector<double>> tot(2, std::vector<double>(nm, 0.0));
    e->tuning = true;
    for (uint32_t round = 0; round <= substeps; ++round) {
        for (int ci = 0; ci < 2; ++ci) {
            for (ModelBuf& m : e->models) m.threads = kCandidates[ci];
            const double dt = e->substeps == 0 ? 0.0 : e->dt_default;
            const int rc = ckmpm_substep(e, dt, e->dt_default);
            if (rc != CKMPM_OK) {
                e->tuning = false;
                return rc;
            }
            CK_HIP(e, hipStreamSynchronize(e->stream));
            if (round == 0) continue;
            for (size_t i = 0; i < nm; ++i) {
                float ms = 0.0f;
                CK_HIP(e, hipEventElapsedTime(&ms, e->models[i].tune_ev[0], e->models[i].tune_ev[1]));
                tot[ci][i] += ms;
            }
        }
    }
    e->tuning = false;
    for (size_t i = 0; i < nm; ++i) e->models[i].threads = tot[1][i] < tot[0][i] ? kCandidates[1] : kCandidates[0];
    return ckmpm_synchronize(e);
}

int ckmpm_set_tuning(ckmpm_engine* e, int32_t model, int32_t threads) {
    if (!e || model < 0 || model >= static_cast<int32_t>(e->models.size())) return CKMPM_ERR_INVALID;
    if (threads != 256 && threads != 512) return Fail(e, CKMPM_ERR_INVALID, "G2P2G workgroup size must be 256 or 512");
    if (e->models[model].threads != threads) ++e->args_version;  // captured graphs hold the old size
    e->models[model].threads = threads;
    return CKMPM_OK;
}

int ckmpm_get_tuning(ckmpm_engine* e, int32_t model, int32_t* threads) {
    if (!e || !threads || model < 0 || model >= static_cast<int32_t>(e->models.size())) return CKMPM_ERR_INVALID;
    *threads = e->models[model].threads;
    return CKMPM_OK;
}

int ckmpm_synchronize(ckmpm_engine* e) {
    if (!e) return CKMPM_ERR_INVALID;
    Status st;
    LayoutCounts lc;
    CK_TRY(ReadStatus(e, st, lc));
    const int code = FlagsToCode(st.flags);
    if (code != CKMPM_OK) return Fail(e, code, "device flags 0x" + std::to_string(st.flags));
    return CKMPM_OK;
}

int ckmpm_get_counters(ckmpm_engine* e, ckmpm_counters* out) {
    if (!e || !out) return CKMPM_ERR_INVALID;
    Status st;
    LayoutCounts lc;
    CK_TRY(ReadStatus(e, st, lc));
    out->active_blocks = lc.A;
    out->neighbor_blocks = lc.Nb;
    out->exterior_blocks = lc.Ex;
    out->flags = st.flags;
    // the slot of the last grid update (substep count - 1; both slots are zero before any)
    out->max_speed = std::sqrt(__builtin_bit_cast(float, st.vmax2_bits[(e->substeps + 1u) & 1u]));
    out->p2g_global_deposits = st.p2g_global;
    out->substeps = e->substeps;
    return CKMPM_OK;
}

int ckmpm_clear_flags(ckmpm_engine* e) {
    if (!e) return CKMPM_ERR_INVALID;
    CK_HIP(e, hipMemsetAsync(&e->status->flags, 0, 2 * sizeof(uint32_t), e->stream));  // flags, p2g_global
    CK_HIP(e, hipStreamSynchronize(e->stream));
    return CKMPM_OK;
}

int ckmpm_model_info(ckmpm_engine* e, int32_t model, uint64_t* n, int32_t* nattr) {
    if (!e || model < 0 || model >= static_cast<int32_t>(e->models.size())) return CKMPM_ERR_INVALID;
    if (n) *n = e->models[model].n;
    if (nattr) *nattr = e->models[model].nattr;
    return CKMPM_OK;
}

int ckmpm_get_positions(ckmpm_engine* e, int32_t model, float* xyz, uint64_t capacity, uint64_t* n) {
    CK_TRY(CheckReady(e));
    if (model < 0 || model >= static_cast<int32_t>(e->models.size()) || !xyz) return CKMPM_ERR_INVALID;
    ModelBuf& m = e->models[model];
    if (capacity < m.n) return Fail(e, CKMPM_ERR_INVALID, "capacity below particle count");
    // staging buffer for the largest model, allocated on the first call and kept (a frame
    // loop calls this once per frame)
    if (!e->pos_stage) {
        uint32_t nmax = 0;
        for (const ModelBuf& q : e->models) nmax = std::max(nmax, q.n);
        CK_TRY(Alloc(e, &e->pos_stage, 3 * static_cast<size_t>(nmax)));
    }
    const LayoutBuf& P = e->L[1 - e->cur];
    LaunchExportPositions(e->dom, P.keys, P.lc, m.off[1 - e->sc], m.st[e->sc], m.n, e->pos_stage,
                          GridFor(e->a_hint + 8), e->stream);
    CK_TRY(CopyOnStream(e, xyz, e->pos_stage, 3 * sizeof(float) * static_cast<size_t>(m.n), hipMemcpyDeviceToHost));
    if (n) *n = m.n;
    return CKMPM_OK;
}

int ckmpm_export_layout(ckmpm_engine* e, int32_t which, uint32_t* keys, uint32_t capacity, uint32_t counts[3]) {
    CK_TRY(CheckReady(e));
    const LayoutBuf& L = e->L[which ? 1 - e->cur : e->cur];
    LayoutCounts lc;
    CK_HIP(e, hipMemcpyAsync(&lc, L.lc, sizeof(lc), hipMemcpyDeviceToHost, e->stream));
    CK_HIP(e, hipStreamSynchronize(e->stream));
    if (counts) {
        counts[0] = lc.A;
        counts[1] = lc.Nb;
        counts[2] = lc.Ex;
    }
    if (keys) {
        if (capacity < lc.Ex) return Fail(e, CKMPM_ERR_INVALID, "capacity below exterior count");
        CK_TRY(CopyOnStream(e, keys, L.keys, lc.Ex * sizeof(uint32_t), hipMemcpyDeviceToHost));
    }
    return CKMPM_OK;
}

int ckmpm_export_ordering(ckmpm_engine* e, int32_t model, int32_t which, uint32_t* perm, uint32_t* offsets,
                          uint32_t offsets_capacity) {
    CK_TRY(CheckReady(e));
    if (model < 0 || model >= static_cast<int32_t>(e->models.size())) return CKMPM_ERR_INVALID;
    ModelBuf& m = e->models[model];
    const int li = which ? 1 - e->cur : e->cur;
    const int oi = which ? 1 - e->sc : e->sc;
    LayoutCounts lc;
    CK_HIP(e, hipMemcpyAsync(&lc, e->L[li].lc, sizeof(lc), hipMemcpyDeviceToHost, e->stream));
    CK_HIP(e, hipStreamSynchronize(e->stream));
    if (perm) CK_TRY(CopyOnStream(e, perm, m.perm[oi], m.n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (offsets) {
        if (offsets_capacity < lc.A + 1) return Fail(e, CKMPM_ERR_INVALID, "offsets capacity below A + 1");
        CK_TRY(CopyOnStream(e, offsets, m.off[oi], (lc.A + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost));
    }
    return CKMPM_OK;
}

int ckmpm_export_state(ckmpm_engine* e, int32_t model, float* soa) {
    CK_TRY(CheckReady(e));
    if (model < 0 || model >= static_cast<int32_t>(e->models.size()) || !soa) return CKMPM_ERR_INVALID;
    ModelBuf& m = e->models[model];
    CK_TRY(CopyOnStream(e, soa, m.st[e->sc], static_cast<size_t>(m.n) * m.nattr * sizeof(float), hipMemcpyDeviceToHost));
    return CKMPM_OK;
}

int ckmpm_import_state(ckmpm_engine* e, int32_t model, const float* soa) {
    CK_TRY(CheckReady(e));
    if (model < 0 || model >= static_cast<int32_t>(e->models.size()) || !soa) return CKMPM_ERR_INVALID;
    ModelBuf& m = e->models[model];
    CK_TRY(CopyOnStream(e, m.st[e->sc], soa, static_cast<size_t>(m.n) * m.nattr * sizeof(float), hipMemcpyHostToDevice));
    return CKMPM_OK;
}

int ckmpm_export_grid_m(ckmpm_engine* e, float* grid, uint64_t capacity_blocks) {
    CK_TRY(CheckReady(e));
    LayoutCounts lc;
    CK_HIP(e, hipMemcpyAsync(&lc, e->L[1 - e->cur].lc, sizeof(lc), hipMemcpyDeviceToHost, e->stream));
    CK_HIP(e, hipStreamSynchronize(e->stream));
    if (capacity_blocks < lc.Nb || !grid) return Fail(e, CKMPM_ERR_INVALID, "capacity below neighbour count");
    CK_TRY(CopyOnStream(e, grid, e->M[e->mprev], static_cast<size_t>(lc.Nb) * kMFloats * sizeof(float), hipMemcpyDeviceToHost));
    return CKMPM_OK;
}

int ckmpm_export_grid_v(ckmpm_engine* e, float* grid, uint64_t capacity_blocks) {
    CK_TRY(CheckReady(e));
    LayoutCounts lc;
    CK_HIP(e, hipMemcpyAsync(&lc, e->L[1 - e->cur].lc, sizeof(lc), hipMemcpyDeviceToHost, e->stream));
    CK_HIP(e, hipStreamSynchronize(e->stream));
    if (capacity_blocks < lc.Nb || !grid) return Fail(e, CKMPM_ERR_INVALID, "capacity below neighbour count");
    CK_TRY(CopyOnStream(e, grid, e->V, static_cast<size_t>(lc.Nb) * kVFloats * sizeof(float), hipMemcpyDeviceToHost));
    return CKMPM_OK;
}

int ckmpm_set_profiling(ckmpm_engine* e, int32_t enable) {
    if (!e) return CKMPM_ERR_INVALID;
    CK_HIP(e, hipStreamSynchronize(e->stream));
    e->profile = enable != 0;
    e->ev_used = 0;
    return CKMPM_OK;
}

int ckmpm_get_phase_times(ckmpm_engine* e, double ms[3], uint64_t* substeps) {
    if (!e || !ms) return CKMPM_ERR_INVALID;
    CK_HIP(e, hipStreamSynchronize(e->stream));
    ms[0] = ms[1] = ms[2] = 0.0;
    const size_t n = e->ev_used / 4;
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            float t = 0.0f;
            CK_HIP(e, hipEventElapsedTime(&t, e->events[4 * i + k], e->events[4 * i + k + 1]));
            ms[k] += t;
        }
    }
    if (substeps) *substeps = n;
    e->ev_used = 0;
    return CKMPM_OK;
}

int ckmpm_destroy(ckmpm_engine* e) {
    if (!e) return CKMPM_ERR_INVALID;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    AdjointDestroy(e);
    ConservationDestroy(e);
    for (hipEvent_t ev : e->events) (void)hipEventDestroy(ev);
    for (ModelBuf& m : e->models)
        for (hipEvent_t ev : m.tune_ev)
            if (ev) (void)hipEventDestroy(ev);
    for (void* p : e->allocs) (void)hipFree(p);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return CKMPM_OK;
}

const char* ckmpm_last_error(const ckmpm_engine* e) { return e ? e->err.c_str() : "null engine"; }

}  // extern "C"

// ============================================================== batched scenes
// Many independent engines advanced together.  Each engine keeps its own stream; the
// graph mode captures two substeps of every engine (the buffer roles ping-pong every
// substep, so two substeps return each engine to the roles it was captured in) as one
// hipGraph forked over the engine streams, and replays it: one launch per two
// substeps of the whole batch instead of ~30 launches per engine per substep.
struct BatchGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<int> roles;         // engine buffer roles at capture
    std::vector<uint32_t> hints;    // block-count hints the launch sizes were chosen for
    std::vector<uint64_t> versions; // engine argument versions at capture
    void* tables = nullptr;         // fused mode: argument tables and workgroup starts
};

struct ckmpm_batch {
    std::vector<ckmpm_engine*> eng;
    hipStream_t cs = nullptr;  // capture and replay stream
    hipEvent_t fork = nullptr, done = nullptr;
    std::vector<hipEvent_t> join;
    // the engines step together, so their buffer roles take one of two configurations
    // (even / odd substep count): one graph each
    BatchGraph g[2];
    // mode 4: one graph per parity of fused launches (each kernel of a substep once for
    // all scenes)
    BatchGraph fz[2];
    BatchGraph fz1[2];  // fused mode, one substep: the odd remainder of a run
    // fused mode: streams and events that run a substep's G2P2G launches side by side
    std::vector<hipStream_t> aux;
    std::vector<hipEvent_t> aux_done;
    bool g2p2g_serial = false;  // fused mode: the G2P2G launches of a substep one after the other
    // mode 3: one graph per engine and parity, replayed on the engine's own stream
    std::vector<BatchGraph> pe[2];
    uint64_t captures = 0;
    std::string err;
};

namespace {

int BatchFail(ckmpm_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    return code;
}

#define BATCH_HIP(b, call)                                                                                   \
    do {                                                                                                     \
        hipError_t _r = (call);                                                                              \
        if (_r != hipSuccess) return BatchFail((b), CKMPM_ERR_HIP, std::string(#call ": ") + hipGetErrorString(_r)); \
    } while (0)

// buffer roles, and the grid-update instantiation (default or custom boundary), that a
// captured graph bakes in
int EngineRoles(const ckmpm_engine* e) {
    return e->cur | (e->sc << 1) | (e->mprev << 2) | (e->bc_custom << 3) | static_cast<int>((e->substeps & 1u) << 4);
}

void EagerSubstep(ckmpm_engine* e) {
    const double dt = e->substeps == 0 ? 0.0 : e->dt_default;
    PhaseGrid(e, dt);
    PhaseParticles(e, dt, e->dt_default);
}

// The graph's replays have completed (its tables are freed).
void ReleaseGraph(BatchGraph& g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    if (g.tables) (void)hipFree(g.tables);
    g.exec = nullptr;
    g.graph = nullptr;
    g.tables = nullptr;
}

int Parity(const ckmpm_batch* b) { return static_cast<int>(b->eng[0]->substeps & 1u); }

bool GraphCurrent(const ckmpm_batch* b, const BatchGraph& g) {
    if (!g.exec) return false;
    for (size_t i = 0; i < b->eng.size(); ++i) {
        const ckmpm_engine* e = b->eng[i];
        if (EngineRoles(e) != g.roles[i] || e->args_version != g.versions[i]) return false;
        // launch sizes follow the block count; recapture once it drifts by more than 1/4
        const uint32_t h = g.hints[i], a = e->a_hint;
        if (a > h + h / 4 + 8 || a + a / 4 + 8 < h) return false;
    }
    return true;
}

int Capture(ckmpm_batch* b, BatchGraph& g) {
    ReleaseGraph(g);
    g.roles.assign(b->eng.size(), -1);
    g.hints.assign(b->eng.size(), 0);
    g.versions.assign(b->eng.size(), 0);
    std::vector<bool> prof(b->eng.size());
    BATCH_HIP(b, hipStreamBeginCapture(b->cs, hipStreamCaptureModeRelaxed));
    (void)hipEventRecord(b->fork, b->cs);
    for (size_t i = 0; i < b->eng.size(); ++i) {
        ckmpm_engine* e = b->eng[i];
        prof[i] = e->profile;
        e->profile = false;
        (void)hipStreamWaitEvent(e->stream, b->fork, 0);
        g.roles[i] = EngineRoles(e);
        g.hints[i] = e->a_hint;
        g.versions[i] = e->args_version;
        for (int k = 0; k < 2; ++k) {
            PhaseGrid(e, e->dt_default);
            PhaseParticles(e, e->dt_default, e->dt_default);
        }
        e->substeps -= 2;  // recorded, not run
        (void)hipEventRecord(b->join[i], e->stream);
        (void)hipStreamWaitEvent(b->cs, b->join[i], 0);
    }
    const hipError_t rc = hipStreamEndCapture(b->cs, &g.graph);
    for (size_t i = 0; i < b->eng.size(); ++i) b->eng[i]->profile = prof[i];
    if (rc != hipSuccess) return BatchFail(b, CKMPM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(rc));
    BATCH_HIP(b, hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
    ++b->captures;
    return CKMPM_OK;
}

// Two substeps of one engine captured on its own stream.
int CaptureEngine(ckmpm_batch* b, ckmpm_engine* e, BatchGraph& g) {
    ReleaseGraph(g);
    g.roles.assign(1, EngineRoles(e));
    g.hints.assign(1, e->a_hint);
    g.versions.assign(1, e->args_version);
    const bool prof = e->profile;
    e->profile = false;
    BATCH_HIP(b, hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed));
    for (int k = 0; k < 2; ++k) {
        PhaseGrid(e, e->dt_default);
        PhaseParticles(e, e->dt_default, e->dt_default);
    }
    e->substeps -= 2;
    const hipError_t rc = hipStreamEndCapture(e->stream, &g.graph);
    e->profile = prof;
    if (rc != hipSuccess) return BatchFail(b, CKMPM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(rc));
    BATCH_HIP(b, hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
    ++b->captures;
    return CKMPM_OK;
}

// Two substeps of every engine as fused launches: each (kernel, variant) of a substep
// runs once for all scenes, its workgroups split by scene (LaunchFused), from argument
// tables built here on the host and copied once.  The stages run in order on the capture
// stream, so every engine's launches keep their order (the G2P2G launches of a substep,
// which are independent, run side by side); the scenes' buffers are disjoint.
// Captures nsub (1 or 2) substeps of every engine as fused launches.
int CaptureFused(ckmpm_batch* b, BatchGraph& g, int nsub) {
    ReleaseGraph(g);
    const size_t ne = b->eng.size();
    g.roles.assign(ne, -1);
    g.hints.assign(ne, 0);
    g.versions.assign(ne, 0);
    FusedRec rec[2];
    rec[0].shared = rec[1].shared = ne > 1;
    for (size_t i = 0; i < ne; ++i) {
        g.roles[i] = EngineRoles(b->eng[i]);
        g.hints[i] = b->eng[i]->a_hint;
        g.versions[i] = b->eng[i]->args_version;
    }
    for (int k = 0; k < nsub; ++k)
        for (ckmpm_engine* e : b->eng) {
            const bool prof = e->profile;
            e->profile = false;
            PhaseGrid(e, e->dt_default, &rec[k]);
            PhaseParticles(e, e->dt_default, e->dt_default, &rec[k]);
            e->profile = prof;
        }
    for (ckmpm_engine* e : b->eng) e->substeps -= nsub;  // recorded, not run
    // recording flipped the buffer roles nsub times; flip them back
    if (nsub & 1)
        for (ckmpm_engine* e : b->eng) FlipRoles(e);
    // host image of the tables: per launch, the argument structs then the workgroup starts
    struct Launch {
        int stage, kernel, variant;
        size_t args, begin;
        uint32_t n, total;
    };
    std::vector<Launch> launches;
    std::vector<uint8_t> img;
    auto place = [&](const void* src, size_t bytes) {
        const size_t off = (img.size() + 255) & ~static_cast<size_t>(255);
        img.resize(off + bytes);
        std::memcpy(img.data() + off, src, bytes);
        return off;
    };
    for (int k = 0; k < nsub; ++k) {
        std::stable_sort(rec[k].ph.begin(), rec[k].ph.end(), [](const FusedPhase& x, const FusedPhase& y) {
            if (x.stage != y.stage) return x.stage < y.stage;
            if (x.kernel != y.kernel) return x.kernel < y.kernel;
            return x.variant < y.variant;
        });
        for (const FusedPhase& p : rec[k].ph) {
            std::vector<uint32_t> begin(p.grids.size() + 1, 0u);
            uint64_t tot = 0;
            for (size_t i = 0; i < p.grids.size(); ++i) {
                begin[i] = static_cast<uint32_t>(tot);
                tot += p.grids[i];
            }
            if (tot > 0x7FFFFFFFull / 1024u) return BatchFail(b, CKMPM_ERR_INVALID, "fused launch too large");
            begin.back() = static_cast<uint32_t>(tot);
            if (p.args.size() != FusedArgsSize(p.kernel) * p.grids.size())
                return BatchFail(b, CKMPM_ERR_INVALID, "fused launch: argument table size");
            Launch l{p.stage, p.kernel, p.variant, 0, 0, static_cast<uint32_t>(p.grids.size()),
                     static_cast<uint32_t>(tot)};
            l.args = place(p.args.data(), p.args.size());
            l.begin = place(begin.data(), begin.size() * sizeof(uint32_t));
            launches.push_back(l);
        }
    }
    BATCH_HIP(b, hipMalloc(&g.tables, img.size()));
    BATCH_HIP(b, hipMemcpy(g.tables, img.data(), img.size(), hipMemcpyHostToDevice));
    uint8_t* dev = static_cast<uint8_t*>(g.tables);
    // the G2P2G launches of a substep (one per kind and workgroup size) are independent:
    // they run side by side on auxiliary streams, so their tails overlap
    size_t most = 0;
    for (size_t j = 0; j < launches.size();) {
        size_t k = j;
        while (k < launches.size() && launches[k].kernel == kFkG2P2G) ++k;
        most = std::max(most, k - j);
        j = k > j ? k : j + 1;
    }
    while (b->aux.size() + 1 < std::max<size_t>(most, 2)) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        BATCH_HIP(b, hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        b->aux.push_back(st);
        BATCH_HIP(b, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        b->aux_done.push_back(ev);
    }
    auto launch = [&](const Launch& l, hipStream_t st) {
        LaunchFused(l.kernel, l.variant, dev + l.args, reinterpret_cast<const uint32_t*>(dev + l.begin), l.n, l.total,
                    st);
    };
    BATCH_HIP(b, hipStreamBeginCapture(b->cs, hipStreamCaptureModeRelaxed));
    for (size_t j = 0; j < launches.size();) {
        size_t k = j;
        while (k < launches.size() && launches[k].kernel == kFkG2P2G) ++k;
        if (k - j < 2 || b->g2p2g_serial) {
            const Launch& l = launches[j];
            LaunchFused(l.kernel, l.variant, dev + l.args, reinterpret_cast<const uint32_t*>(dev + l.begin), l.n,
                        l.total, b->cs);
            ++j;
            continue;
        }
        (void)hipEventRecord(b->fork, b->cs);
        for (size_t q = j; q < k; ++q) {
            const Launch& l = launches[q];
            hipStream_t st = b->cs;
            if (q > j) {
                st = b->aux[q - j - 1];
                (void)hipStreamWaitEvent(st, b->fork, 0);
            }
            LaunchFused(l.kernel, l.variant, dev + l.args, reinterpret_cast<const uint32_t*>(dev + l.begin), l.n,
                        l.total, st);
        }
        for (size_t q = j + 1; q < k; ++q) {
            (void)hipEventRecord(b->aux_done[q - j - 1], b->aux[q - j - 1]);
            (void)hipStreamWaitEvent(b->cs, b->aux_done[q - j - 1], 0);
        }
        j = k;
    }
    const hipError_t rc = hipStreamEndCapture(b->cs, &g.graph);
    if (rc != hipSuccess) return BatchFail(b, CKMPM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(rc));
    BATCH_HIP(b, hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
    ++b->captures;
    return CKMPM_OK;
}

bool EngineGraphCurrent(const ckmpm_engine* e, const BatchGraph& g) {
    if (!g.exec || EngineRoles(e) != g.roles[0] || e->args_version != g.versions[0]) return false;
    const uint32_t h = g.hints[0], a = e->a_hint;
    return !(a > h + h / 4 + 8 || a + a / 4 + 8 < h);
}

}  // namespace

extern "C" {

int ckmpm_batch_create(ckmpm_engine* const* engines, uint32_t count, ckmpm_batch** out) {
    if (!engines || !out || count == 0) return CKMPM_ERR_INVALID;
    *out = nullptr;
    for (uint32_t i = 0; i < count; ++i) {
        if (!engines[i] || !engines[i]->setup_done || engines[i]->adj) return CKMPM_ERR_INVALID;
        for (uint32_t j = 0; j < i; ++j)
            if (engines[j] == engines[i]) return CKMPM_ERR_INVALID;  // an engine appears once
    }
    ckmpm_batch* b = new ckmpm_batch();
    b->eng.assign(engines, engines + count);
    b->join.resize(count, nullptr);
    bool ok = hipStreamCreateWithFlags(&b->cs, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&b->done, hipEventDisableTiming) == hipSuccess;
    for (uint32_t i = 0; ok && i < count; ++i) ok = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        ckmpm_batch_destroy(b);
        return CKMPM_ERR_HIP;
    }
    *out = b;
    return CKMPM_OK;
}

int ckmpm_batch_run(ckmpm_batch* b, uint32_t substeps, int32_t mode) {
    if (!b || mode < 0 || mode > 4) return CKMPM_ERR_INVALID;
    if (mode == 3) {
        uint32_t left = substeps;
        bool first = false;
        for (ckmpm_engine* e : b->eng) first |= e->substeps == 0;
        if (first && left) {
            for (ckmpm_engine* e : b->eng) EagerSubstep(e);
            --left;
        }
        const uint32_t pairs = left / 2;
        if (pairs) {
            const int par = Parity(b);
            b->pe[par].resize(b->eng.size());
            for (size_t i = 0; i < b->eng.size(); ++i) {
                ckmpm_engine* e = b->eng[i];
                if (!EngineGraphCurrent(e, b->pe[par][i])) {
                    BATCH_HIP(b, hipStreamSynchronize(e->stream));
                    const int rc = CaptureEngine(b, e, b->pe[par][i]);
                    if (rc != CKMPM_OK) return rc;
                }
            }
            for (uint32_t p = 0; p < pairs; ++p)
                for (size_t i = 0; i < b->eng.size(); ++i)
                    BATCH_HIP(b, hipGraphLaunch(b->pe[par][i].exec, b->eng[i]->stream));
            for (ckmpm_engine* e : b->eng) e->substeps += 2ull * pairs;
        }
        if (left & 1u)
            for (ckmpm_engine* e : b->eng) EagerSubstep(e);
        const hipError_t r = hipGetLastError();
        if (r != hipSuccess) return BatchFail(b, CKMPM_ERR_HIP, std::string("launch: ") + hipGetErrorString(r));
        return CKMPM_OK;
    }
    if (mode == 0) {
        // one scene after the other, each with plain per-substep launches
        for (ckmpm_engine* e : b->eng) {
            for (uint32_t k = 0; k < substeps; ++k) EagerSubstep(e);
            const int rc = ckmpm_synchronize(e);
            if (rc != CKMPM_OK) return BatchFail(b, rc, "engine: " + e->err);
        }
        return CKMPM_OK;
    }
    if (mode == 1) {
        for (uint32_t k = 0; k < substeps; ++k)
            for (ckmpm_engine* e : b->eng) EagerSubstep(e);
        return CKMPM_OK;
    }
    // graph modes: a pending first substep (dt = 0) and an odd remainder run eagerly
    uint32_t left = substeps;
    bool first = false;
    for (ckmpm_engine* e : b->eng) first |= e->substeps == 0;
    if (first && left) {
        // every engine advances one substep (dt = 0 for those that had not started)
        for (ckmpm_engine* e : b->eng) EagerSubstep(e);
        --left;
    }
    const uint32_t pairs = left / 2;
    // fused mode runs an odd remainder as a one-substep fused graph; graph mode eagerly
    const bool single = mode == 4 && (left & 1u);
    if (pairs || single) {
        // the replay stream runs after everything already queued on the engine streams
        for (size_t i = 0; i < b->eng.size(); ++i) {
            BATCH_HIP(b, hipEventRecord(b->join[i], b->eng[i]->stream));
            BATCH_HIP(b, hipStreamWaitEvent(b->cs, b->join[i], 0));
        }
    }
    for (int nsub = 2; nsub >= 1; --nsub) {
        const uint32_t reps = nsub == 2 ? pairs : (single ? 1u : 0u);
        if (!reps) continue;
        const int par = Parity(b);
        BatchGraph& g = mode == 4 ? (nsub == 2 ? b->fz[par] : b->fz1[par]) : b->g[par];
        if (!GraphCurrent(b, g)) {
            BATCH_HIP(b, hipStreamSynchronize(b->cs));
            for (ckmpm_engine* e : b->eng) BATCH_HIP(b, hipStreamSynchronize(e->stream));
            const int rc = mode == 4 ? CaptureFused(b, g, nsub) : Capture(b, g);
            if (rc != CKMPM_OK) return rc;
        }
        for (uint32_t p = 0; p < reps; ++p) BATCH_HIP(b, hipGraphLaunch(g.exec, b->cs));
        for (ckmpm_engine* e : b->eng) {
            e->substeps += static_cast<uint64_t>(nsub) * reps;
            if (nsub == 1) FlipRoles(e);  // the engines' buffers now play the other roles
        }
    }
    if (pairs || single) {
        // later work on an engine stream waits for the replays
        BATCH_HIP(b, hipEventRecord(b->done, b->cs));
        for (ckmpm_engine* e : b->eng) BATCH_HIP(b, hipStreamWaitEvent(e->stream, b->done, 0));
    }
    if ((left & 1u) && !single)
        for (ckmpm_engine* e : b->eng) EagerSubstep(e);
    const hipError_t r = hipGetLastError();
    if (r != hipSuccess) return BatchFail(b, CKMPM_ERR_HIP, std::string("launch: ") + hipGetErrorString(r));
    return CKMPM_OK;
}

int ckmpm_batch_autotune(ckmpm_batch* b, uint32_t pairs) {
    if (!b) return CKMPM_ERR_INVALID;
    if (pairs == 0) return CKMPM_OK;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    BATCH_HIP(b, hipEventCreate(&t0));
    BATCH_HIP(b, hipEventCreate(&t1));
    // device time of `pairs` replays of the two-substep fused graph, after one untimed
    // replay (which captures the graph for the candidate)
    auto timed = [&](double* ms) -> int {
        int rc = ckmpm_batch_run(b, 2, 4);
        if (rc != CKMPM_OK) return rc;
        if (hipEventRecord(t0, b->cs) != hipSuccess) return BatchFail(b, CKMPM_ERR_HIP, "hipEventRecord");
        rc = ckmpm_batch_run(b, 2 * pairs, 4);
        if (rc != CKMPM_OK) return rc;
        float f = 0.0f;
        if (hipEventRecord(t1, b->cs) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
            hipEventElapsedTime(&f, t0, t1) != hipSuccess)
            return BatchFail(b, CKMPM_ERR_HIP, "autotune timing");
        *ms = f;
        return CKMPM_OK;
    };
    auto set_kind = [&](int kind, int threads) {
        for (ckmpm_engine* e : b->eng)
            for (size_t i = 0; i < e->models.size(); ++i)
                if (e->models[i].kind == kind) (void)ckmpm_set_tuning(e, static_cast<int32_t>(i), threads);
    };
    auto set_serial = [&](bool serial) { (void)ckmpm_batch_set_serial(b, serial ? 1 : 0); };
    std::vector<int> kinds;
    for (ckmpm_engine* e : b->eng)
        for (const ModelBuf& m : e->models)
            if (std::find(kinds.begin(), kinds.end(), m.kind) == kinds.end()) kinds.push_back(m.kind);
    int rc = CKMPM_OK;
    // one coordinate at a time: each kind's workgroup size, then the G2P2G launch order
    for (int kind : kinds) {
        double best = 0.0;
        int best_threads = 256;
        for (int threads : {256, 512}) {
            set_kind(kind, threads);
            double ms = 0.0;
            if ((rc = timed(&ms)) != CKMPM_OK) break;
            if (threads == 256 || ms < best) {
                best = ms;
                best_threads = threads;
            }
        }
        if (rc != CKMPM_OK) break;
        set_kind(kind, best_threads);
    }
    if (rc == CKMPM_OK) {
        double side = 0.0, serial = 0.0;
        set_serial(false);
        rc = timed(&side);
        if (rc == CKMPM_OK) {
            set_serial(true);
            rc = timed(&serial);
        }
        if (rc == CKMPM_OK) set_serial(serial < side);
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (rc != CKMPM_OK) return rc;
    return ckmpm_batch_synchronize(b);
}

int ckmpm_batch_get_serial(const ckmpm_batch* b, int32_t* serial) {
    if (!b || !serial) return CKMPM_ERR_INVALID;
    *serial = b->g2p2g_serial ? 1 : 0;
    return CKMPM_OK;
}

int ckmpm_batch_set_serial(ckmpm_batch* b, int32_t serial) {
    if (!b) return CKMPM_ERR_INVALID;
    if (b->g2p2g_serial != (serial != 0)) {
        BATCH_HIP(b, hipStreamSynchronize(b->cs));  // the graphs are rebuilt for the new order
        b->g2p2g_serial = serial != 0;
        for (int p = 0; p < 2; ++p) {
            ReleaseGraph(b->fz[p]);
            ReleaseGraph(b->fz1[p]);
        }
    }
    return CKMPM_OK;
}

int ckmpm_batch_synchronize(ckmpm_batch* b) {
    if (!b) return CKMPM_ERR_INVALID;
    BATCH_HIP(b, hipStreamSynchronize(b->cs));
    for (ckmpm_engine* e : b->eng) {
        const int rc = ckmpm_synchronize(e);
        if (rc != CKMPM_OK) return BatchFail(b, rc, "engine: " + e->err);
    }
    return CKMPM_OK;
}

int ckmpm_batch_captures(const ckmpm_batch* b, uint64_t* captures) {
    if (!b || !captures) return CKMPM_ERR_INVALID;
    *captures = b->captures;
    return CKMPM_OK;
}

int ckmpm_batch_destroy(ckmpm_batch* b) {
    if (!b) return CKMPM_ERR_INVALID;
    if (b->cs) (void)hipStreamSynchronize(b->cs);
    ReleaseGraph(b->g[0]);
    ReleaseGraph(b->g[1]);
    ReleaseGraph(b->fz[0]);
    ReleaseGraph(b->fz[1]);
    ReleaseGraph(b->fz1[0]);
    ReleaseGraph(b->fz1[1]);
    for (int p = 0; p < 2; ++p)
        for (BatchGraph& g : b->pe[p]) ReleaseGraph(g);
    for (hipEvent_t ev : b->join)
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : b->aux_done) (void)hipEventDestroy(ev);
    for (hipStream_t st : b->aux) (void)hipStreamDestroy(st);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->cs) (void)hipStreamDestroy(b->cs);
    delete b;
    return CKMPM_OK;
}

const char* ckmpm_batch_last_error(const ckmpm_batch* b) { return b ? b->err.c_str() : "null batch"; }

}  // extern "C"
