/*
 * ckmpm.h - C-ABI of the MI355X-native Compact-Kernel MPM substep (libckmpm.so).
 *
 * The reference exposes no C-ABI: its API is a header-only C++ template engine
 * (include/mpm_engine.cuh) whose configuration is compile-time.  Each entry point
 * below names the reference interface it replaces (file:line, relative to the
 * reference tree).  All pointers are host pointers; sizes are element counts.
 * Every call returns CKMPM_OK or an error code; ckmpm_last_error() gives the text.
 * One engine is single-threaded (the caller serialises calls on it); any number of
 * engines may coexist in a process (the reference engine is a singleton,
 * mpm_engine.cuh:80-84).  No call terminates the process (the reference calls
 * exit()/abort(), mpm_engine.cuh:545-549,739-744,908).
 */
#ifndef CKMPM_H
#define CKMPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ckmpm_engine ckmpm_engine;

/* status codes (the reference: cudaCheckError -> exit, cuda_util.cuh:9-22; abort on
 * capacity, mpm_engine.cuh:250-254,307-312,739-744,809-816,858-865) */
enum {
    CKMPM_OK = 0,
    CKMPM_ERR_INVALID = 1,        /* bad argument or call order */
    CKMPM_ERR_HIP = 2,            /* HIP runtime error */
    CKMPM_ERR_CAPACITY = 3,       /* block list exceeded block_capacity */
    CKMPM_ERR_NONFINITE = 4,      /* NaN / inf grid velocity */
    CKMPM_ERR_OUT_OF_WINDOW = 5,  /* a particle moved farther than its P2G window */
    CKMPM_ERR_OUT_OF_DOMAIN = 6,  /* a particle block left the domain */
    CKMPM_ERR_NO_DEVICE = 7       /* no gfx950 device / library built without HIP */
};

/* device error flag bits reported in ckmpm_counters.flags */
#define CKMPM_FLAG_OUT_OF_WINDOW 0x1u
#define CKMPM_FLAG_OUT_OF_DOMAIN 0x2u
#define CKMPM_FLAG_NONFINITE 0x4u
#define CKMPM_FLAG_CAPACITY 0x8u
#define CKMPM_FLAG_MISSING_BLOCK 0x10u

/* constitutive models: mpm_material.cuh:10-18 (FixedCorotated, DruckerPragerStvkhencky,
 * Fluid, VonMises, NonAssociatedCamClay) plus Neo-Hookean, which the reference lacks */
enum {
    CKMPM_FIXED_COROTATED = 0,
    CKMPM_NEO_HOOKEAN = 1,
    CKMPM_DRUCKER_PRAGER = 2,
    CKMPM_FLUID = 3,
    CKMPM_VON_MISES = 4,
    CKMPM_CAM_CLAY = 5 /* NonAssociatedCamClay */
};

/* MPMDomain<MPMDomainRange<Dx,Dy,Dz>, MPMDomainOffset<Ox,Oy,Oz>> (mpm_domain.h:9-61):
 * extent and offset in 4^3-cell blocks; each extent 1..1024 (up to 2^30 blocks: the two
 * dense block tables of an engine then take 8 GiB) */
typedef struct {
    int32_t dim[3];
    int32_t offset[3];
} ckmpm_domain_desc;

/* MPMConfigBase getters (mpm_config.h:17-123) as runtime fields */
typedef struct {
    float dx;                /* GetDx */
    float gravity;           /* GetGravity (applied along y) */
    float mass_clamp;        /* GetMassClamp */
    float cfl;               /* GetCfl */
    float dt;                /* GetDt (default / maximum dt) */
    uint32_t fps;            /* GetFps */
    uint32_t block_capacity; /* GetMaxActiveBlockCount */
} ckmpm_sim_desc;

/* MPMMaterial<Model>::parameter_ (mpm_material.cuh:62-175) */
typedef struct {
    int32_t kind;             /* CKMPM_FIXED_COROTATED ... */
    float lambda, mu;         /* FC, NH, DP */
    float bulk, gamma;        /* fluid: pressure = bulk (J^-gamma - 1) */
    float viscosity, J0;      /* fluid */
    float friction_angle_deg; /* DP; the reference hard-codes 30 (mpm_material.cuh:369) */
    float yield_stress;       /* von Mises (VonMisesMaterialParameter, mpm_material.cuh:236-253) */
    /* Cam-clay (NonAssociatedCamClayMaterialParameter, mpm_material.cuh:189-214): initial
     * logJp, beta (p_min = -beta p0), hardening factor xi, slope M, hardening on/off */
    float alpha0, beta, hardening_factor, m;
    int32_t hardening_on;
} ckmpm_material_desc;

typedef struct {
    uint32_t active_blocks;   /* particle blocks (activePartitionBlockCount_) */
    uint32_t neighbor_blocks; /* grid blocks (activePartitionNeighborBlockCount_) */
    uint32_t exterior_blocks; /* activePartitionExteriorBlockCount_ */
    uint32_t flags;           /* CKMPM_FLAG_* accumulated since the last clear */
    float max_speed;          /* sqrt(max |v|^2) of the last grid update, over the nodes
                                 holding at least 2^-20 of the lightest particle's mass */
    uint32_t p2g_global_deposits; /* particles whose P2G momentum went to the grid with FP32
                                     global atomics (deposit above the fixed-point range of
                                     the launch) since the last clear; 0 in steady runs */
    uint64_t substeps;
} ckmpm_counters;

/* MPMEngine::GetInstance + Initialize (mpm_engine.cuh:80-84,170-191) */
int ckmpm_create(const ckmpm_domain_desc* domain, const ckmpm_sim_desc* sim, ckmpm_engine** out);

/* MPMEngine::InitializeParticle for one MPMModel (mpm_engine.cuh:91-168,
 * mpm_model.h:19-33).  vel holds nvel = 1 or n velocities.  lambda_p / mu_p are
 * optional per-particle Lame parameters of a solid (both NULL: the material's); they
 * become two extra state attributes that move with the particles. */
int ckmpm_add_model(ckmpm_engine* eng, const ckmpm_material_desc* material, float mass, float volume,
                    const float* xyz, uint64_t n, const float* vel, uint64_t nvel, const float* lambda_p,
                    const float* mu_p, int32_t* model_index);

/* MPMEngine::InitialSetup (mpm_engine.cuh:215-383): binning, block activation,
 * rasterize; then the first grid update with dt = 0. */
int ckmpm_initial_setup(ckmpm_engine* eng);

/* One iteration of the Simulate substep loop (mpm_engine.cuh:519-876) with
 * caller-chosen dt (advection) and next_dt (stress deposit): grid update with dt,
 * G2P2G, rebin.  Enqueued on the engine's stream; returns without a host sync.
 * The reference's first iteration has dt = 0 (mpm_engine.cuh:479).  The engine's
 * simulated time then advances by next_dt, as the frame loop's timestamp does
 * (mpm_engine.cuh:870-871). */
int ckmpm_substep(ckmpm_engine* eng, double dt, double next_dt);

/* Simulate's frame loop with the reference dt rules (mpm_engine.cuh:514-876):
 * next_dt = min(default dt, time to frame end, dx cfl / max|v|); stops on a
 * non-finite velocity.  Runs `frames` frames. */
int ckmpm_run_frames(ckmpm_engine* eng, uint32_t frames);

/* Grid boundary conditions.  The reference's grid update zeroes the velocity of every grid
 * node within 2 blocks (8 cells) of a domain face (mpm_algorithm.cuh:759-762,800-805)
 * unless the scene's config supplies ProcessGridCellVelocity(cell, velocity, frame),
 * which then replaces that rule (:795-799).  The scenes' hooks are regions and walls:
 * a fixed box and slip walls (tests/fire_hydrant/test_fire_hydrant.cuh:148-185), slip
 * walls (tests/fluid_with_two_longs/test_fluid_with_two_longs.cuh:133-161), rotating
 * clamps and sticky walls (tests/twisting_bar/test_twisting_bar.cuh:126-168).  This
 * description expresses them as data.  For every grid node with mass above the clamp,
 * after v = p / m + g dt: the regions are applied in order to the node's cell (integer
 * cell coordinates, both sub-grids alike, as the hook receives them), then the walls. */
enum {
    CKMPM_WALLS_STICKY = 0, /* v = 0 within 8 cells of a face (the reference's default) */
    CKMPM_WALLS_SLIP = 1,   /* there, remove the component along n / |n|, n = sum of the
                               inward normals of the faces within 8 cells */
    CKMPM_WALLS_NONE = 2
};
enum {
    CKMPM_REGION_FIXED = 0,    /* v = 0 */
    CKMPM_REGION_VELOCITY = 1, /* v = value */
    CKMPM_REGION_ROTATION = 2  /* v = value x (cell dx - center): rotation with angular velocity value */
};
#define CKMPM_MAX_REGIONS 16
typedef struct {
    int32_t lo[3], hi[3]; /* cells lo <= c < hi */
    int32_t kind;         /* CKMPM_REGION_* */
    float value[3];
    float center[3];
} ckmpm_region;
typedef struct {
    int32_t walls; /* CKMPM_WALLS_* */
    uint32_t nregions;
    ckmpm_region regions[CKMPM_MAX_REGIONS];
} ckmpm_boundary_desc;
/* Replaces the boundary description (NULL: the reference's default, sticky walls and no
 * regions).  Takes effect from the next substep, also in captured batch graphs; not to
 * be called while the engine's stream is being captured.  An engine with the adjoint
 * enabled keeps the default (its grid backward differentiates that rule only). */
int ckmpm_set_boundary(ckmpm_engine* eng, const ckmpm_boundary_desc* boundary);

/* Rigid particles: replaces the scene hooks GetExistRigidParticle, GetRigidParticleCount
 * and GetRigidParticleVelocity(frame) (mpm_config.h:92-106) and their use in G2P2G
 * (mpm_algorithm.cuh:1065-1068 flag, 1316-1328 and 1609-1624 update).  The first `count`
 * particles of a solid model (input order) are rigid: their G2P takes nothing from the
 * grid (C = 0, v = 0, so F is unchanged) and they move by dt * velocity; they still
 * deposit their mass and stress in P2G.  `count` is fixed before ckmpm_initial_setup
 * (the flag travels with each particle as one more state attribute); `velocity` (m/s)
 * may be changed between substeps, as the reference's velocity of the frame. */
int ckmpm_set_rigid(ckmpm_engine* eng, int32_t model, uint64_t count, const float velocity[3]);

/* Waits for the stream; returns the first error encoded in the device flags. */
int ckmpm_synchronize(ckmpm_engine* eng);
int ckmpm_get_counters(ckmpm_engine* eng, ckmpm_counters* out);
int ckmpm_clear_flags(ckmpm_engine* eng);

/* RetrieveParticleBuffer + ExportParticleToModel (mpm_algorithm.cuh:1977-2020,
 * mpm_engine.cuh:385-454): global FP32 positions of one model, AoS xyz, storage order. */
int ckmpm_get_positions(ckmpm_engine* eng, int32_t model, float* xyz, uint64_t capacity, uint64_t* n);

/* --- state export / import, in the engine's storage layout.  Used by parity tests
 * and by callers that checkpoint; no reference counterpart (its buffers are only
 * reachable through the debug printers, mpm_particle.cuh:96-200). */
int ckmpm_model_info(ckmpm_engine* eng, int32_t model, uint64_t* n, int32_t* nattr);
/* which = 0: current layout, 1: previous layout (the one the grid mass/momentum and
 * the stored particle state were produced in) */
int ckmpm_export_layout(ckmpm_engine* eng, int32_t which, uint32_t* keys, uint32_t capacity, uint32_t counts[3]);
/* Ordering of the current (which = 0) or previous (1) layout: perm[t] = slot | move << 27
 * (slot of the stored state read by processing position t; move = (dx+1)*9+(dy+1)*3+(dz+1)
 * of the particle's block since the state was written), offsets[j] = first position of
 * block j (A + 1 entries). */
int ckmpm_export_ordering(ckmpm_engine* eng, int32_t model, int32_t which, uint32_t* perm, uint32_t* offsets,
                          uint32_t offsets_capacity);
/* SoA [attr][n] float32, storage order: u (3, position relative to the origin 4 b dx of
 * the block b that wrote it: block j of the previous ordering for slot t in its range),
 * then E = F - I (9, row-major) or q = J - 1 (fluid), then logJp (DP); strains are
 * stored as deviations from the identity so that FP32 keeps their relative precision */
int ckmpm_export_state(ckmpm_engine* eng, int32_t model, float* soa);
int ckmpm_import_state(ckmpm_engine* eng, int32_t model, const float* soa);
/* grid mass/momentum of the previous layout, [block][m,px,py,pz][node 128] */
int ckmpm_export_grid_m(ckmpm_engine* eng, float* grid, uint64_t capacity_blocks);
/* grid velocity of the current layout, [block][vx,vy,vz][node 128] */
int ckmpm_export_grid_v(ckmpm_engine* eng, float* grid, uint64_t capacity_blocks);

/* Simulate(config, path, collectConservationMetric = true) (mpm_engine.cuh:494-603;
 * kernel CollectConservationMetric, mpm_algorithm.cuh:1083-1357): after the grid update
 * of every ckmpm_substep / ckmpm_run_frames substep, per model, the G2P velocity v_p of
 * each particle on the updated grid at its position x_p, summed into the Lagrangian
 * linear momentum sum v_p and angular momentum sum x_p x v_p (per unit mass, as the
 * reference).  Recorded on the device without a host sync (the reference syncs each
 * substep); batched runs (ckmpm_batch_run) record nothing.  capacity = substeps kept; 0
 * turns recording off.  Enabling clears the log; a substep that finds it full fails with
 * CKMPM_ERR_CAPACITY before it starts. */
int ckmpm_set_conservation(ckmpm_engine* eng, uint32_t capacity);
/* The model's records, n of them, in the layout of the reference's
 * conservationMetric_model_<i>.bin (mpm_engine.cuh:597-602): 13 floats per substep =
 * simulated time, Eulerian linear and angular momentum (zero: the reference never fills
 * them), Lagrangian linear and angular momentum.  records holds capacity substeps
 * (13 floats each); capacity < n is an error; records = NULL only sets n.  Waits for
 * the engine's stream. */
int ckmpm_get_conservation(ckmpm_engine* eng, int32_t model, float* records, uint64_t capacity, uint64_t* n);

/* Per-phase device time, measured with hipEvents on the engine's stream, like the
 * reference's MPMBenchmark registry (mpm_benchmark.h:12-79; phases "G2P2G (Cuda)",
 * "Activate Neighbor", ... at mpm_engine.cuh:485-492).  ms[0] grid update, ms[1]
 * G2P2G (all models), ms[2] rebin and block activation, summed over the substeps
 * since the last call; the call waits for the stream and resets the sums. */
int ckmpm_set_profiling(ckmpm_engine* eng, int32_t enable);
int ckmpm_get_phase_times(ckmpm_engine* eng, double ms[3], uint64_t* substeps);

/* Tuning (no reference counterpart: its launch shapes are compile-time constants).
 * The G2P2G workgroup size of each model is 256 or 512 threads (one particle block per
 * workgroup).  ckmpm_autotune runs 1 + `substeps` rounds of one real substep per
 * candidate, interleaved (the scene advances at sim.dt, the first with dt = 0); the
 * first round is not timed.  Per model it keeps the size with the least G2P2G device
 * time: 2 (substeps + 1) substeps in all. */
int ckmpm_autotune(ckmpm_engine* eng, uint32_t substeps);
int ckmpm_set_tuning(ckmpm_engine* eng, int32_t model, int32_t threads);
int ckmpm_get_tuning(ckmpm_engine* eng, int32_t model, int32_t* threads);

/* MPMEngine destructor (the reference frees almost nothing) */
int ckmpm_destroy(ckmpm_engine* eng);
const char* ckmpm_last_error(const ckmpm_engine* eng);

/* ---------------------------------------------------------------- adjoint
 * No reference counterpart (the reference has no differentiation).  Reverse-mode
 * substeps for fixed-corotated and Neo-Hookean models with per-particle lambda and mu
 * (gradients w.r.t. those and the per-particle initial velocity).  Usage:
 *   ckmpm_adjoint_enable (after ckmpm_add_model, before ckmpm_initial_setup);
 *   forward: ckmpm_checkpoint_save every S substeps, ckmpm_substep;
 *   backward: ckmpm_adjoint_reset, ckmpm_adjoint_seed (dL/dx at the last state), then
 *   per segment from the last: ckmpm_checkpoint_restore, ckmpm_adjoint_record(1), the
 *   segment's ckmpm_substep calls again, ckmpm_adjoint_record(0), one
 *   ckmpm_adjoint_substep per recorded substep; finally ckmpm_adjoint_gradients. */
typedef struct ckmpm_checkpoint ckmpm_checkpoint;
/* tape of up to `segment` substeps; also keeps particle identities (input order) */
int ckmpm_adjoint_enable(ckmpm_engine* eng, uint32_t segment);
int ckmpm_checkpoint_save(ckmpm_engine* eng, ckmpm_checkpoint** out);
int ckmpm_checkpoint_restore(ckmpm_engine* eng, const ckmpm_checkpoint* ck);
/* the checkpoint's memory returns to its engine for the next ckmpm_checkpoint_save
 * (freed with the engine; a checkpoint outliving its engine only releases its handle) */
int ckmpm_checkpoint_free(ckmpm_checkpoint* ck);
/* No counterpart in the reference, which builds a new engine for every run: puts an
 * adjoint-enabled engine with one model back to its state right after
 * ckmpm_initial_setup (same positions, setup order and block layout) with new initial
 * velocities (nvel = 1 or n, 3 floats each) and new per-particle Lame parameters (n
 * each, input order), the grid mass/momentum rasterized again and the substep count at
 * 0.  Allocates nothing, so a gradient loop on one scene reuses one engine. */
int ckmpm_adjoint_rewind(ckmpm_engine* eng, const float* vel, uint64_t nvel, const float* lambda_p,
                         const float* mu_p);
/* on != 0: the following substeps are recorded (the tape restarts empty).  The engine
 * keeps two tapes of `segment` substeps (the second allocated on first use): a
 * recording goes to the tape not being undone, so the replay of one segment runs beside
 * the undo of the next one (ckmpm_adjoint_substep runs on the engine's second stream). */
int ckmpm_adjoint_record(ckmpm_engine* eng, int32_t on);
/* zero every particle and grid adjoint */
int ckmpm_adjoint_reset(ckmpm_engine* eng);
/* dL/dx of the model's particles at the current state, input order, 3 floats each */
int ckmpm_adjoint_seed(ckmpm_engine* eng, int32_t model, const float* xbar);
/* back-propagate through the last substep of the tape recorded last and drop it from
 * the tape; queued on the engine's second stream after the recording (asynchronous: the
 * calls that read or clear the adjoints wait for it) */
int ckmpm_adjoint_substep(ckmpm_engine* eng);
/* after the first substep has been undone: dL/dlambda, dL/dmu (n each) and dL/dv0
 * (3 n), input order; any pointer may be null */
int ckmpm_adjoint_gradients(ckmpm_engine* eng, int32_t model, float* lambda_bar, float* mu_bar, float* v0_bar);
/* positions (3 per particle) in input order; needs the adjoint enabled */
int ckmpm_get_positions_input_order(ckmpm_engine* eng, int32_t model, float* xyz);

/* ---------------------------------------------------------------- batched scenes
 * No reference counterpart (the reference engine is a singleton, mpm_engine.cuh:80-84):
 * many independent engines advanced together.  The batch does not own the engines;
 * each must have run ckmpm_initial_setup, and appears in the batch once
 * (CKMPM_ERR_INVALID otherwise).  Every engine uses its own sim.dt (an engine that has
 * not run a substep yet starts with dt = 0, as mpm_engine.cuh:479 does). */
typedef struct ckmpm_batch ckmpm_batch;
int ckmpm_batch_create(ckmpm_engine* const* engines, uint32_t count, ckmpm_batch** out);
/* mode 0: one engine after the other, plain launches, synchronised per engine (the
 *         sequential single-scene baseline);
 * mode 1: substep k of every engine enqueued on the engines' own streams, no sync;
 * mode 2: as 1, but pairs of substeps replay one hipGraph captured over all engine
 *         streams (recaptured when an engine's block count drifts by more than 1/4);
 * mode 3: pairs of substeps replay one hipGraph per engine on the engine's own stream;
 * mode 4: as 2, but each kernel of a substep is one launch for all engines (one per
 *         material kind and workgroup size for G2P2G, per boundary kind for the grid
 *         update), its workgroups split by engine; about 9 launches per substep for the
 *         whole batch (grid update, G2P2G per kind, rebin count, scan, scatter with the
 *         extension mark, emit).  The graph reads argument tables that it holds in device memory. */
int ckmpm_batch_run(ckmpm_batch* b, uint32_t substeps, int32_t mode);
/* waits for every engine; returns the first engine error */
int ckmpm_batch_synchronize(ckmpm_batch* b);
int ckmpm_batch_captures(const ckmpm_batch* b, uint64_t* captures);
/* Tuning of mode 4 (no reference counterpart): per material kind in the batch, the G2P2G
 * workgroup size (256 or 512 threads) of every model of that kind, then whether a
 * substep's G2P2G launches (one per kind and size) run side by side or one after the
 * other.  Each candidate is timed with HIP events over `pairs` replays of the two-substep
 * fused graph after one untimed replay; the engines advance 2 (pairs + 1) substeps per
 * candidate (2 kinds + 2 candidates in all).  The choice stays with the batch and its
 * engines (ckmpm_get_tuning, ckmpm_batch_get_serial). */
int ckmpm_batch_autotune(ckmpm_batch* b, uint32_t pairs);
int ckmpm_batch_get_serial(const ckmpm_batch* b, int32_t* serial);
int ckmpm_batch_set_serial(ckmpm_batch* b, int32_t serial);
int ckmpm_batch_destroy(ckmpm_batch* b);
const char* ckmpm_batch_last_error(const ckmpm_batch* b);

/* ---------------------------------------------------------------- sparse block table
 * Replaces BucketedCuckooHashTable<Pair, 16, MurmurHash3_32> (gpu_hash_table.cuh:274-552):
 * linear block key -> uint32 value.  Same hash (MurmurHash3_x86_32 of the int3 block
 * coordinate, :39-118), same seeds (std::mt19937(2), :303-309), same bucket choice
 * (hash % buckets) and Find protocol (:463-513).  The build is deterministic: the table
 * contents are a function of the (key, value) set alone (the reference's atomic cuckoo
 * insert, :378-461, is not).  Keys must be unique and inside the domain. */
typedef struct ckmpm_hash ckmpm_hash;
typedef struct {
    uint32_t buckets;            /* ceil(max_keys / (16 * 0.7)) */
    uint32_t slots_per_bucket;   /* 16 */
    uint32_t stash_capacity;
    uint32_t keys;               /* pairs of the last build */
    uint32_t occupied_slots;     /* pairs in buckets */
    uint32_t stash;              /* pairs in the stash */
    uint32_t seeds[3];
    uint32_t fill_histogram[17]; /* buckets holding 0..16 pairs */
} ckmpm_hash_info_t;
/* Allocate (gpu_hash_table.cuh:343-354) for at most max_keys pairs of the domain's keys */
int ckmpm_hash_create(const ckmpm_domain_desc* domain, uint32_t max_keys, ckmpm_hash** out);
/* Build from host arrays (replaces a pass of CooperativeInsert, gpu_hash_table.cuh:378-461) */
int ckmpm_hash_build(ckmpm_hash* h, const uint32_t* keys, const uint32_t* values, uint32_t n);
/* Same from device arrays; the table's stream is synchronised before return */
int ckmpm_hash_build_device(ckmpm_hash* h, const uint32_t* keys, const uint32_t* values, uint32_t n);
/* The build of ckmpm_hash_build_device, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The caller passes the
 * device arrays: keys and values (n pairs), and the table to fill, sized for max_keys
 * pairs as ckmpm_hash_create sizes it: slots of ceil(10 max_keys / 112) * 16 words, stash
 * of 256 + max_keys / 64 words, and one word for the stash count.  The result is the
 * table ckmpm_hash_create(domain, max_keys) + ckmpm_hash_build_device give, word for word.
 * The temporaries come from a static device scratch of the library (11.5 MB on each GPU
 * it runs on, held until the process ends), so max_keys is at most
 * CKMPM_HASH_NOALLOC_MAX_KEYS (CKMPM_ERR_CAPACITY above it, or for n > max_keys).  Enqueued
 * on stream (a hipStream_t; null: the null stream), which is synchronised before return.
 * Every call on one GPU shares that GPU's scratch: calls that overlap on one GPU race. */
#define CKMPM_HASH_NOALLOC_MAX_KEYS (1u << 18)
int ckmpm_hash_build_noalloc(const ckmpm_domain_desc* domain, uint32_t max_keys, const uint32_t* keys,
                             const uint32_t* values, uint32_t n, uint64_t* slots, uint64_t* stash, uint32_t* nstash,
                             void* stream);
/* Find (CooperativeFind, gpu_hash_table.cuh:463-513) of host keys; missing -> 0xFFFFFFFF */
int ckmpm_hash_find(ckmpm_hash* h, const uint32_t* keys, uint32_t n, uint32_t* values);
/* Same on device arrays, enqueued on the table's stream (no sync).  cooperative = 1: a
 * 16-lane quarter wavefront per probe (the reference's tile of 16); 0: one thread per key;
 * 2: LDS-probed, for tables of at most 2560 buckets (created for up to 28.6k keys;
 * CKMPM_ERR_INVALID above): each workgroup stages the table's keys in LDS once, then
 * probes there. */
int ckmpm_hash_find_device(ckmpm_hash* h, const uint32_t* keys, uint32_t n, uint32_t* values, int32_t cooperative);
int ckmpm_hash_synchronize(ckmpm_hash* h);
int ckmpm_hash_info(ckmpm_hash* h, ckmpm_hash_info_t* info);
/* Raw table: buckets*16 slot words (key << 32 | value, empty = ~0) and the stash */
int ckmpm_hash_export(ckmpm_hash* h, uint64_t* slots, uint64_t slots_capacity, uint64_t* stash,
                      uint32_t stash_capacity, uint32_t* nstash);
int ckmpm_hash_destroy(ckmpm_hash* h);
const char* ckmpm_hash_last_error(const ckmpm_hash* h);

#ifdef __cplusplus
}
#endif

#endif /* CKMPM_H */
