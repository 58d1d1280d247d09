// Internal layout of an engine, shared by engine.hip and adjoint.hip (not installed).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/ckmpm.h"
#include "kernels.h"

namespace ckmpm {

struct ModelBuf {
    MaterialDev mat{};
    int kind = 0;
    int nattr = 0;
    uint32_t n = 0;
    float* st[2] = {nullptr, nullptr};
    uint32_t* perm[2] = {nullptr, nullptr};
    uint32_t* off[2] = {nullptr, nullptr};
    uint32_t* code = nullptr;
    uint32_t* cnt = nullptr;
    uint32_t* base = nullptr;
    uint32_t* cnt_e = nullptr;
    uint32_t* eoff = nullptr;
    float* blkmax = nullptr;  // [cap] largest P2G deposit bound per particle block
    int pp_attr = 0;      // state attribute of per-particle (lambda, mu); 0: none
    int rigid_attr = 0;   // state attribute of the rigid flag (1: rigid); 0: none
    uint64_t rigid_count = 0;
    float rigid_v[3] = {0.f, 0.f, 0.f};
    int threads = 256;    // G2P2G workgroup size (256 or 512; ckmpm_autotune chooses)
    hipEvent_t tune_ev[2] = {nullptr, nullptr};  // around the G2P2G launch while tuning
    // setup inputs (host), released after the initial setup
    std::vector<float> x0, v0, lam0, mu0;
    uint32_t nvel = 0;
};

struct LayoutBuf {
    uint32_t* keys = nullptr;
    uint32_t* table = nullptr;
    LayoutCounts* lc = nullptr;
};

inline int AttrCount(int kind) {
    return (kind == kDruckerPrager || kind == kCamClay) ? 13 : (kind == kFluid ? 4 : 12);
}



}  // namespace ckmpm

using namespace ckmpm;

struct AdjointState;  // adjoint.hip
struct ConservationLog;  // conservation.hip

struct ckmpm_engine {
    DomainDev dom{};
    ckmpm_sim_desc sim{};
    uint32_t cap = 0;
    hipStream_t stream = nullptr;
    LayoutBuf L[2];
    int cur = 0;          // current layout; 1 - cur is the previous one
    float* M[2] = {nullptr, nullptr};
    int mprev = 0;        // M[mprev] holds the previous layout's mass / momentum
    float* V = nullptr;
    Status* status = nullptr;
    BoundaryDev* bc = nullptr;   // grid boundary description (custom = 0: the default rule)
    bool bc_custom = false;      // host copy of bc->custom
    uint32_t* bitmap = nullptr;       // setup: one bit per domain block
    uint32_t nwords = 0;
    uint32_t* word_off = nullptr;
    uint32_t* ext_map = nullptr;      // layout extension: two bits per domain block
    uint32_t ext_tiles = 0;           // map tiles of ext_tile_words words
    uint32_t ext_tile_words = 0;      // ExtTileWords of the domain
    uint32_t ext_tile_shift = 0;      // block key -> map tile shift (ExtTileShift)
    int emit_variant = 0;             // EmitVariant of the domain
    uint32_t* ext_cnt = nullptr;      // [ext_tiles][2] neighbour / exterior blocks marked per tile
    uint32_t* tsum = nullptr;         // rebin count tile sums
    uint32_t* scan_tmp = nullptr;
    uint32_t* total = nullptr;
    float* pos_stage = nullptr;  // ckmpm_get_positions staging (3 floats per particle of the largest model)
    std::vector<ModelBuf> models;
    int sc = 0;           // state / ordering buffer index of the current step
    bool setup_done = false;
    uint32_t a_hint = 0;  // particle-block count at the last host read
    uint64_t substeps = 0;
    // bumped when a value that kernel arguments carry changes after setup (the rigid
    // velocity): captured batch graphs hold arguments by value and are captured again
    uint64_t args_version = 0;
    double dt_default = 0;
    double time = 0, cur_dt = 0;
    std::string err;
    std::vector<void*> allocs;
    // phase timing: 4 events per substep (before grid, after grid, after G2P2G, after rebin)
    bool profile = false;
    bool tuning = false;  // record tune_ev around each model's G2P2G launch
    std::vector<hipEvent_t> events;
    size_t ev_used = 0;
    // adjoint support (null unless ckmpm_adjoint_enable was called)
    AdjointState* adj = nullptr;
    // per-substep conservation metric (null unless ckmpm_set_conservation enabled it)
    ConservationLog* cons = nullptr;
};

namespace ckmpm {

inline int Fail(ckmpm_engine* e, int code, const std::string& msg) {
    if (e) e->err = msg;
    return code;
}

#define CK_HIP(e, call)                                                                                  \
    do {                                                                                                 \
        hipError_t _r = (call);                                                                          \
        if (_r != hipSuccess) return Fail((e), CKMPM_ERR_HIP, std::string(#call ": ") + hipGetErrorString(_r)); \
    } while (0)


// A stream for work that must run beside the engine stream (the adjoint's undo beside
// the replay).  The runtime shares a few hardware queues among the process's streams,
// and two streams on one queue run one after the other; a stream created with a CU mask
// (here every CU) gets a queue of its own.  Such a queue is not free: while one exists,
// the 8M-particle substep's graph replays ran 7% slower (0.709 -> 0.757 ms), so it is
// used only where the overlap pays for that.  Falls back to a shared stream.
inline hipError_t CreateOwnQueueStream(hipStream_t* s) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0) {
        std::vector<uint32_t> mask((static_cast<size_t>(cus) + 31) / 32, 0xFFFFFFFFu);
        if (hipExtStreamCreateWithCUMask(s, static_cast<uint32_t>(mask.size()), mask.data()) == hipSuccess)
            return hipSuccess;
        (void)hipGetLastError();
    }
    return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}

template <typename T>
inline int Alloc(ckmpm_engine* e, T** p, size_t count) {
    void* q = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t r = hipMalloc(&q, bytes);
    if (r != hipSuccess) return Fail(e, CKMPM_ERR_HIP, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(r));
    e->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return CKMPM_OK;
}

// adjoint hooks (adjoint.hip); no-ops unless ckmpm_adjoint_enable was called
void AdjointOnSetupOrder(ckmpm_engine* e, size_t model, const uint32_t* order, const float* xyz);
int AdjointBeforeSubstep(ckmpm_engine* e, double dt, double nd);
void AdjointAfterSubstep(ckmpm_engine* e);
// after a successful initial setup: keep its state for ckmpm_adjoint_rewind
void AdjointAfterSetup(ckmpm_engine* e);
// error when the substep tape is recording and full (checked before a substep starts)
int AdjointCheckTape(ckmpm_engine* e);
// identity arrays of a model for the next substep (null, null without differentiation)
void AdjointIdBuffers(ckmpm_engine* e, size_t model, const uint32_t** in, uint32_t** out);
void AdjointDestroy(ckmpm_engine* e);

// conservation metric hooks (conservation.hip); no-ops unless ckmpm_set_conservation
// enabled it.  Check before a substep starts (error when the log is full); record after
// its grid update, at the simulated time e->time.
int ConservationCheck(ckmpm_engine* e);
void ConservationAfterGrid(ckmpm_engine* e);
void ConservationDestroy(ckmpm_engine* e);

#define CK_TRY(x)                   \
    do {                            \
        int _rc = (x);              \
        if (_rc != CKMPM_OK) return _rc; \
    } while (0)

}  // namespace ckmpm
