// Per-substep conservation metric (the reference's Simulate(..., collectConservationMetric
// = true), mpm_engine.cuh:494-603, kernel CollectConservationMetric,
// mpm_algorithm.cuh:1083-1357).
//
// After the grid update of a substep, each particle's G2P velocity on the updated grid,
// v_p = 1/2 sum_g sum_i w_ig v_ig at its position x_p (the state the substep's G2P2G is
// about to read), is summed per model: linear sum v_p and angular sum x_p x v_p, per unit
// mass as in the reference.  The reference's two Eulerian entries are never written
// there (mpm_algorithm.cuh:1096-1099: only momentum[2] and [3] receive atomics) and stay
// zero here too.  Records are kept on the device, one per substep and model (FP64
// sums), with the simulated time on the host; the caller reads them in the reference's
// file layout: 13 floats per substep (time, Eulerian linear, Eulerian angular,
// Lagrangian linear, Lagrangian angular).
#include <algorithm>
#include <string>
#include <vector>

#include "engine_impl.h"
#include "stencil.h"

struct ConservationLog {
    uint32_t capacity = 0;     // substeps
    uint32_t len = 0;          // recorded substeps
    double* dev = nullptr;     // [capacity][models][6]: linear xyz, angular xyz
    std::vector<double> time;  // simulated time of each record
};

namespace ckmpm {

namespace {

constexpr int kCT = 256;

struct ConsArgs {
    DomainDev dom;
    const uint32_t* keys;
    const uint32_t* table;
    const LayoutCounts* lc;
    const float* V;
    const float* st;  // state read by the coming G2P2G (slot order)
    const uint32_t* perm;
    const uint32_t* off;
    uint32_t n;
    double* out;      // 6 sums of this record and model
};

__device__ __forceinline__ double WaveSum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// One workgroup per particle block (grid-stride): the particles of the block in the
// current processing order, positions relative to the block as G2P2G forms them, the 16
// grid velocities read through the table (the 8 blocks b + {0,1}^3 of the current
// layout's neighbour set hold every stencil node).
__global__ __launch_bounds__(kCT) void ConservationKernel(ConsArgs a) {
    __shared__ double part[kCT / 64][6];
    const uint32_t A = a.lc->A, Nb = a.lc->Nb;
    const size_t N = a.n;
    const float dx = a.dom.dx, inv_dx = a.dom.inv_dx;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t j = blockIdx.x; j < A; j += gridDim.x) {
        int bb[3];
        CoordsOf(a.dom, a.keys[j], bb[0], bb[1], bb[2]);
        for (uint32_t t = a.off[j] + threadIdx.x; t < a.off[j + 1]; t += kCT) {
            const uint32_t pe = a.perm[t];
            const uint32_t s = pe & kSlotMask;
            const int mv = static_cast<int>(pe >> kMoveShift);
            if (s >= a.n || mv >= kDirections) continue;  // flagged by G2P2G
            const int mov[3] = {mv / 9 - 1, (mv / 3) % 3 - 1, mv % 3 - 1};
            float x[3], xs[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                x[k] = __fsub_rn(a.st[k * N + s], __fmul_rn(static_cast<float>(4 * mov[k]), dx));
                xs[k] = __fmul_rn(x[k], inv_dx);
            }
            Stencil32 sp[2];
            MakeStencilPair(xs, sp[0], sp[1]);
            float v[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int g = 0; g < 2; ++g)
#pragma unroll
                for (int o = 0; o < 8; ++o) {
                    const int d[3] = {(o >> 2) & 1, (o >> 1) & 1, o & 1};
                    int nb[3], loc[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const int c = sp[g].c[k] + d[k];  // node cell relative to block j
                        nb[k] = bb[k] + (c >> 2);         // arithmetic shift: floor division
                        loc[k] = c & 3;
                    }
                    if (!InDomain(a.dom, nb[0], nb[1], nb[2])) continue;
                    const uint32_t bi = a.table[KeyOf(a.dom, nb[0], nb[1], nb[2])];
                    if (bi == kEmpty || bi >= Nb) continue;
                    const float w = NodeWeight(sp[g], d[0], d[1], d[2]);
                    const float* vb = a.V + static_cast<size_t>(bi) * kVFloats + g * 64 + loc[0] * 16 + loc[1] * 4 + loc[2];
#pragma unroll
                    for (int k = 0; k < 3; ++k) v[k] += w * vb[k * kNodesPerBlock];
                }
            double p[3], vd[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                vd[k] = 0.5 * static_cast<double>(v[k]);
                // global position as ckmpm_get_positions forms it
                p[k] = static_cast<double>(__fadd_rn(__fmul_rn(static_cast<float>(4 * bb[k]), dx), x[k]));
            }
            acc[0] += vd[0];
            acc[1] += vd[1];
            acc[2] += vd[2];
            acc[3] += p[1] * vd[2] - p[2] * vd[1];
            acc[4] += p[2] * vd[0] - p[0] * vd[2];
            acc[5] += p[0] * vd[1] - p[1] * vd[0];
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        const double r = WaveSum(acc[c]);
        if (lane == 0) part[wv][c] = r;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        double r = 0.0;
#pragma unroll
        for (int w = 0; w < kCT / 64; ++w) r += part[w][threadIdx.x];
        if (r != 0.0) atomicAdd(&a.out[threadIdx.x], r);
    }
}

}  // namespace

int ConservationCheck(ckmpm_engine* e) {
    const ConservationLog* c = e->cons;
    if (c && c->len >= c->capacity)
        return Fail(e, CKMPM_ERR_CAPACITY,
                    "conservation log full (" + std::to_string(c->capacity) +
                        " substeps): read it and call ckmpm_set_conservation again");
    return CKMPM_OK;
}

void ConservationAfterGrid(ckmpm_engine* e) {
    ConservationLog* c = e->cons;
    if (!c || c->len >= c->capacity) return;
    const LayoutBuf& C = e->L[e->cur];
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(e->a_hint + e->a_hint / 4 + 8, 8192));
    for (size_t mi = 0; mi < e->models.size(); ++mi) {
        const ModelBuf& m = e->models[mi];
        ConsArgs a{};
        a.dom = e->dom;
        a.keys = C.keys;
        a.table = C.table;
        a.lc = C.lc;
        a.V = e->V;
        a.st = m.st[e->sc];
        a.perm = m.perm[e->sc];
        a.off = m.off[e->sc];
        a.n = m.n;
        a.out = c->dev + (static_cast<size_t>(c->len) * e->models.size() + mi) * 6;
        ConservationKernel<<<grid, kCT, 0, e->stream>>>(a);
    }
    c->time.push_back(e->time);
    c->len++;
}

void ConservationDestroy(ckmpm_engine* e) {
    if (!e->cons) return;
    if (e->cons->dev) (void)hipFree(e->cons->dev);
    delete e->cons;
    e->cons = nullptr;
}

}  // namespace ckmpm

using namespace ckmpm;

extern "C" int ckmpm_set_conservation(ckmpm_engine* e, uint32_t capacity) {
    if (!e) return CKMPM_ERR_INVALID;
    if (!e->setup_done) return Fail(e, CKMPM_ERR_INVALID, "call ckmpm_initial_setup first");
    CK_HIP(e, hipStreamSynchronize(e->stream));
    ConservationDestroy(e);
    if (capacity == 0) return CKMPM_OK;
    ConservationLog* c = new ConservationLog();
    c->capacity = capacity;
    const size_t bytes = static_cast<size_t>(capacity) * e->models.size() * 6 * sizeof(double);
    const hipError_t r = hipMalloc(&c->dev, bytes);
    if (r != hipSuccess) {
        delete c;
        return Fail(e, CKMPM_ERR_HIP, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(r));
    }
    e->cons = c;
    CK_HIP(e, hipMemsetAsync(c->dev, 0, bytes, e->stream));
    CK_HIP(e, hipStreamSynchronize(e->stream));
    c->time.reserve(capacity);
    return CKMPM_OK;
}

extern "C" int ckmpm_get_conservation(ckmpm_engine* e, int32_t model, float* records, uint64_t capacity,
                                      uint64_t* n) {
    if (!e || !n || model < 0 || model >= static_cast<int32_t>(e->models.size())) return CKMPM_ERR_INVALID;
    const ConservationLog* c = e->cons;
    *n = c ? c->len : 0;
    if (!c || c->len == 0 || !records) return CKMPM_OK;  // records = NULL: size query
    if (capacity < c->len) return Fail(e, CKMPM_ERR_INVALID, "records holds fewer than n substeps");
    std::vector<double> host(static_cast<size_t>(c->len) * e->models.size() * 6);
    CK_HIP(e, hipMemcpyAsync(host.data(), c->dev, host.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    CK_HIP(e, hipStreamSynchronize(e->stream));
    for (uint32_t r = 0; r < c->len; ++r) {
        float* o = records + 13 * static_cast<size_t>(r);
        const double* h = host.data() + (static_cast<size_t>(r) * e->models.size() + model) * 6;
        o[0] = static_cast<float>(c->time[r]);
        for (int k = 1; k < 7; ++k) o[k] = 0.0f;  // Eulerian entries: never filled by the reference
        for (int k = 0; k < 6; ++k) o[7 + k] = static_cast<float>(h[k]);
    }
    return CKMPM_OK;
}
