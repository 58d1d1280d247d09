/*
 * ckmpm.hpp - header-only C++17 facade over the C-ABI in ckmpm.h.
 *
 * The reference's API is a compile-time C++ template engine used from scene mains
 * (tests/colliding_sphere/test_colliding_sphere.cu:5-19):
 *     simulator->Initialize(config);
 *     simulator->InitializeParticle(config, modelList);
 *     simulator->InitialSetup(config);
 *     simulator->Simulate(config, ...);
 * This facade keeps those nouns and that call sequence with runtime descriptors:
 * Domain (mpm_domain.h:9-61), Config (mpm_config.h:17-123), Material
 * (mpm_material.cuh:62-175), Model (mpm_model.h:19-33) and Engine (mpm_engine.cuh:28-974).
 * Errors throw ckmpm::Error carrying the CKMPM_ERR_* code and the engine's message (the
 * reference exits the process instead, cuda_util.cuh:9-22).  Engines are move-only
 * owners of their device memory; any number may coexist (no singleton).
 */
#ifndef CKMPM_HPP
#define CKMPM_HPP

#include <array>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "ckmpm.h"

namespace ckmpm {

class Error : public std::runtime_error {
   public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const noexcept { return code_; }

   private:
    int code_;
};

namespace detail {
inline void Check(int rc, const char* call, const char* msg) {
    if (rc == CKMPM_OK) return;
    std::string w = std::string(call) + " failed (" + std::to_string(rc) + ")";
    if (msg && *msg) w += ": " + std::string(msg);
    throw Error(rc, w);
}
}  // namespace detail

/* MPMDomain<MPMDomainRange<..>, MPMDomainOffset<..>>: extent and offset in 4^3-cell blocks */
struct Domain {
    std::array<int32_t, 3> dim{};
    std::array<int32_t, 3> offset{0, 0, 0};
    ckmpm_domain_desc desc() const {
        ckmpm_domain_desc d{};
        for (int k = 0; k < 3; ++k) {
            d.dim[k] = dim[k];
            d.offset[k] = offset[k];
        }
        return d;
    }
};

/* MPMConfigBase getters as fields (GetDx, GetGravity, GetMassClamp, GetCfl, GetDt, GetFps,
 * GetMaxActiveBlockCount) */
struct Config {
    float dx = 0.0f;
    float gravity = 0.0f;
    float mass_clamp = 0.0f;
    float cfl = 0.5f;
    float dt = 1e-4f;
    uint32_t fps = 48;
    uint32_t max_active_block_count = 1u << 16;
    ckmpm_sim_desc desc() const {
        return ckmpm_sim_desc{dx, gravity, mass_clamp, cfl, dt, fps, max_active_block_count};
    }
};

/* ComputeLameParameters (mpm_material.cuh:276-283): {lambda, mu} of Young's modulus E and
 * Poisson ratio nu */
inline std::array<float, 2> LameParameters(double E, double nu) {
    return {static_cast<float>(E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))), static_cast<float>(E / (2.0 * (1.0 + nu)))};
}

/* MPMMaterial<Model>{parameters} */
struct Material {
    ckmpm_material_desc d{};
    static Material FixedCorotated(float lambda, float mu) { return Solid(CKMPM_FIXED_COROTATED, lambda, mu); }
    static Material NeoHookean(float lambda, float mu) { return Solid(CKMPM_NEO_HOOKEAN, lambda, mu); }
    /* the reference fixes the friction angle at 30 degrees (mpm_material.cuh:369) */
    static Material DruckerPrager(float lambda, float mu, float friction_angle_deg = 30.0f) {
        Material m = Solid(CKMPM_DRUCKER_PRAGER, lambda, mu);
        m.d.friction_angle_deg = friction_angle_deg;
        return m;
    }
    /* MPMMaterial<kVonMises>{mu, lambda, yieldStress}: StVK-Hencky with von Mises return */
    static Material VonMises(float lambda, float mu, float yield_stress) {
        Material m = Solid(CKMPM_VON_MISES, lambda, mu);
        m.d.yield_stress = yield_stress;
        return m;
    }
    /* MPMMaterial<kNonAssociatedCamClay>{mu, lambda, alpha0, beta, hardeningFactor, M,
     * hardeningOn} */
    static Material CamClay(float lambda, float mu, float alpha0, float beta, float hardening_factor, float m_slope,
                            bool hardening_on = true) {
        Material m = Solid(CKMPM_CAM_CLAY, lambda, mu);
        m.d.alpha0 = alpha0;
        m.d.beta = beta;
        m.d.hardening_factor = hardening_factor;
        m.d.m = m_slope;
        m.d.hardening_on = hardening_on ? 1 : 0;
        return m;
    }
    /* pressure = bulk (J^-gamma - 1), viscous stress viscosity (C + C^T) */
    static Material Fluid(float bulk, float gamma, float viscosity, float J0 = 1.0f) {
        Material m;
        m.d.kind = CKMPM_FLUID;
        m.d.bulk = bulk;
        m.d.gamma = gamma;
        m.d.viscosity = viscosity;
        m.d.J0 = J0;
        return m;
    }

   private:
    static Material Solid(int kind, float lambda, float mu) {
        Material m;
        m.d.kind = kind;
        m.d.lambda = lambda;
        m.d.mu = mu;
        m.d.J0 = 1.0f;
        m.d.friction_angle_deg = 30.0f;
        return m;
    }
};

/* MPMModel{mass, volume, positions, velocities, material}: positions and velocities are
 * xyz triples; velocities holds one triple (shared) or one per particle.  lambda_p / mu_p
 * (both empty, or one value per particle) override a solid's Lame parameters per
 * particle. */
struct Model {
    float mass = 0.0f;
    float volume = 0.0f;
    std::vector<float> positions;
    std::vector<float> velocities{0.0f, 0.0f, 0.0f};
    Material material;
    std::vector<float> lambda_p, mu_p;
    uint64_t count() const { return positions.size() / 3; }
};

/* The reference's particle assets (the .bin files under asset/particle_data, read by the scene headers,
 * e.g. tests/chimney_ball/test_chimney_ball.cuh:150-160): raw little-endian float32
 * xyz triples, nothing else.  The count is the file size / 12. */
inline std::vector<float> ReadRawPositions(const std::string& path) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw Error(CKMPM_ERR_INVALID, "cannot open " + path);
    long bytes = -1;
    if (std::fseek(f, 0, SEEK_END) == 0) bytes = std::ftell(f);
    bool ok = bytes >= 0 && bytes % 12 == 0 && std::fseek(f, 0, SEEK_SET) == 0;
    std::vector<float> out;
    if (ok) {
        out.resize(static_cast<size_t>(bytes) / 4);
        ok = out.empty() || std::fread(out.data(), 4, out.size(), f) == out.size();
    }
    std::fclose(f);
    if (!ok) throw Error(CKMPM_ERR_INVALID, path + " is not a whole number of float32 xyz triples");
    return out;
}

inline void WriteRawPositions(const std::string& path, const std::vector<float>& xyz) {
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw Error(CKMPM_ERR_INVALID, "cannot create " + path);
    const bool ok = xyz.empty() || std::fwrite(xyz.data(), 4, xyz.size(), f) == xyz.size();
    if (std::fclose(f) != 0 || !ok) throw Error(CKMPM_ERR_INVALID, "cannot write " + path);
}

/* The reference's per-frame particle export (mpm_engine.cuh:446-452: a "position"
 * attribute, written by partio as model_<i>_particle_frame_<f>.bgeo, mpm_io.h:93-131).
 * partio (a git submodule of the reference, not vendored) writes the classic Houdini
 * binary geometry layout, all numbers big-endian: the magic 'Bgeo' (int32), 'V', the
 * version 5 (int32), then eight int32 counts (points, primitives, point groups, primitive
 * groups, point attributes, vertex attributes, primitive attributes, detail attributes;
 * only the points are non-zero here, the position being the point's own P), every point
 * as x y z w with w = 1 (float32), and the bytes 0x00 0xff that close the (empty) extra
 * section.  The file is not compressed (a name ending in .bgeo, not .bgeo.gz). */
inline void WriteBgeoPositions(const std::string& path, const std::vector<float>& xyz) {
    if (xyz.size() % 3 != 0) throw Error(CKMPM_ERR_INVALID, "positions are not xyz triples");
    const uint64_t n = xyz.size() / 3;
    if (n > 0x7FFFFFFFu) throw Error(CKMPM_ERR_INVALID, "too many points for a bgeo file");
    std::vector<unsigned char> buf;
    buf.reserve(41 + 16 * n + 2);
    auto put32 = [&](uint32_t v) {
        for (int s = 24; s >= 0; s -= 8) buf.push_back(static_cast<unsigned char>(v >> s));
    };
    auto putf = [&](float f) {
        uint32_t v;
        std::memcpy(&v, &f, 4);
        put32(v);
    };
    put32(0x4267656Fu);  // 'B' 'g' 'e' 'o'
    buf.push_back('V');
    put32(5u);
    put32(static_cast<uint32_t>(n));
    for (int k = 0; k < 7; ++k) put32(0u);
    for (uint64_t i = 0; i < n; ++i) {
        putf(xyz[3 * i]);
        putf(xyz[3 * i + 1]);
        putf(xyz[3 * i + 2]);
        putf(1.0f);
    }
    buf.push_back(0x00);
    buf.push_back(0xFF);
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw Error(CKMPM_ERR_INVALID, "cannot create " + path);
    const bool ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    if (std::fclose(f) != 0 || !ok) throw Error(CKMPM_ERR_INVALID, "cannot write " + path);
}

/* The scene's ProcessGridCellVelocity hook as data (ckmpm_boundary_desc): regions in
 * order, then walls within 8 cells of the domain faces. */
struct Boundary {
    ckmpm_boundary_desc d{};
    explicit Boundary(int32_t walls = CKMPM_WALLS_STICKY) { d.walls = walls; }
    Boundary& Fixed(std::array<int32_t, 3> lo, std::array<int32_t, 3> hi) {
        return Add(lo, hi, CKMPM_REGION_FIXED, {0, 0, 0}, {0, 0, 0});
    }
    Boundary& Velocity(std::array<int32_t, 3> lo, std::array<int32_t, 3> hi, std::array<float, 3> v) {
        return Add(lo, hi, CKMPM_REGION_VELOCITY, v, {0, 0, 0});
    }
    /* v = omega x (cell dx - center) */
    Boundary& Rotation(std::array<int32_t, 3> lo, std::array<int32_t, 3> hi, std::array<float, 3> omega,
                       std::array<float, 3> center) {
        return Add(lo, hi, CKMPM_REGION_ROTATION, omega, center);
    }

   private:
    Boundary& Add(std::array<int32_t, 3> lo, std::array<int32_t, 3> hi, int32_t kind, std::array<float, 3> value,
                  std::array<float, 3> center) {
        if (d.nregions >= CKMPM_MAX_REGIONS) throw Error(CKMPM_ERR_INVALID, "more than CKMPM_MAX_REGIONS regions");
        ckmpm_region& r = d.regions[d.nregions++];
        for (int k = 0; k < 3; ++k) {
            r.lo[k] = lo[k];
            r.hi[k] = hi[k];
            r.value[k] = value[k];
            r.center[k] = center[k];
        }
        r.kind = kind;
        return *this;
    }
};

/* MPMEngine: Initialize at construction, then InitializeParticle, InitialSetup and
 * Substep / Simulate. */
class Engine {
   public:
    Engine(const Domain& domain, const Config& config) : config_(config) {
        const ckmpm_domain_desc d = domain.desc();
        const ckmpm_sim_desc s = config.desc();
        detail::Check(ckmpm_create(&d, &s, &e_), "ckmpm_create", nullptr);
    }
    ~Engine() {
        if (e_) ckmpm_destroy(e_);
    }
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;
    Engine(Engine&& o) noexcept : e_(std::exchange(o.e_, nullptr)), config_(o.config_) {}
    Engine& operator=(Engine&& o) noexcept {
        if (this != &o) {
            if (e_) ckmpm_destroy(e_);
            e_ = std::exchange(o.e_, nullptr);
            config_ = o.config_;
        }
        return *this;
    }

    /* InitializeParticle for one model (mpm_engine.cuh:91-168); the arrays are copied.
     * Returns the model index. */
    int32_t InitializeParticle(const Model& m) {
        if (m.positions.size() % 3 != 0 || m.velocities.size() % 3 != 0)
            throw Error(CKMPM_ERR_INVALID, "positions and velocities are xyz triples");
        const bool per_particle = !m.lambda_p.empty() || !m.mu_p.empty();
        if (per_particle && (m.lambda_p.size() != m.count() || m.mu_p.size() != m.count()))
            throw Error(CKMPM_ERR_INVALID, "lambda_p and mu_p are both empty or one value per particle");
        int32_t index = -1;
        Call(ckmpm_add_model(e_, &m.material.d, m.mass, m.volume, m.positions.data(), m.count(), m.velocities.data(),
                             m.velocities.size() / 3, m.lambda_p.empty() ? nullptr : m.lambda_p.data(),
                             m.mu_p.empty() ? nullptr : m.mu_p.data(), &index),
             "ckmpm_add_model");
        return index;
    }
    void InitializeParticle(const std::vector<Model>& models) {
        for (const Model& m : models) InitializeParticle(m);
    }
    /* InitialSetup (mpm_engine.cuh:215-383) */
    void InitialSetup() { Call(ckmpm_initial_setup(e_), "ckmpm_initial_setup"); }
    /* one substep; enqueued on the engine's stream, no host sync */
    void Substep(double dt, double next_dt) { Call(ckmpm_substep(e_, dt, next_dt), "ckmpm_substep"); }
    /* `count` substeps at the configured dt; the engine's first substep advects with
     * dt = 0 (as Simulate's first iteration, mpm_engine.cuh:479) */
    void Substeps(uint32_t count) {
        bool first = Counters().substeps == 0;
        for (uint32_t k = 0; k < count; ++k, first = false) Substep(first ? 0.0 : config_.dt, config_.dt);
    }
    /* Simulate's frame loop with the reference dt rules (mpm_engine.cuh:514-876) */
    void Simulate(uint32_t frames) { Call(ckmpm_run_frames(e_, frames), "ckmpm_run_frames"); }
    void Synchronize() { Call(ckmpm_synchronize(e_), "ckmpm_synchronize"); }
    ckmpm_counters Counters() {
        ckmpm_counters c{};
        Call(ckmpm_get_counters(e_, &c), "ckmpm_get_counters");
        return c;
    }
    /* grid boundary (replaces the default sticky walls, as a scene hook does) */
    void SetBoundary(const Boundary& b) { Call(ckmpm_set_boundary(e_, &b.d), "ckmpm_set_boundary"); }
    void ResetBoundary() { Call(ckmpm_set_boundary(e_, nullptr), "ckmpm_set_boundary"); }
    /* per-model G2P2G workgroup size chosen by timing real substeps (the scene advances) */
    void Autotune(uint32_t substeps = 3) { Call(ckmpm_autotune(e_, substeps), "ckmpm_autotune"); }
    /* RetrieveParticleBuffer + ExportParticleToModel: global positions, storage order */
    std::vector<float> Positions(int32_t model) {
        uint64_t n = 0;
        int32_t nattr = 0;
        Call(ckmpm_model_info(e_, model, &n, &nattr), "ckmpm_model_info");
        std::vector<float> xyz(3 * n);
        Call(ckmpm_get_positions(e_, model, xyz.data(), n, &n), "ckmpm_get_positions");
        return xyz;
    }
    /* Simulate(..., collectConservationMetric = true): record the per-substep metric of
     * every model for up to `capacity` substeps (0: off); enabling clears the log */
    void SetConservation(uint32_t capacity) { Call(ckmpm_set_conservation(e_, capacity), "ckmpm_set_conservation"); }
    /* the model's records, 13 floats per substep, as the reference's
     * conservationMetric_model_<i>.bin holds them */
    std::vector<float> Conservation(int32_t model) {
        uint64_t n = 0;
        Call(ckmpm_get_conservation(e_, model, nullptr, 0, &n), "ckmpm_get_conservation");
        std::vector<float> r(13 * n);
        if (n) Call(ckmpm_get_conservation(e_, model, r.data(), n, &n), "ckmpm_get_conservation");
        return r;
    }
    ckmpm_engine* handle() const noexcept { return e_; }

   private:
    void Call(int rc, const char* call) { detail::Check(rc, call, e_ ? ckmpm_last_error(e_) : nullptr); }
    ckmpm_engine* e_ = nullptr;
    Config config_;
};

/* Many independent engines advanced together (ckmpm_batch_*); the engines must outlive
 * the batch. */
class Batch {
   public:
    /* kFused: each kernel of a substep is one launch for all engines (the fastest) */
    enum Mode { kSequential = 0, kStreams = 1, kSingleGraph = 2, kPerSceneGraphs = 3, kFused = 4 };
    explicit Batch(const std::vector<Engine*>& engines) {
        std::vector<ckmpm_engine*> h;
        for (Engine* e : engines) h.push_back(e->handle());
        detail::Check(ckmpm_batch_create(h.data(), static_cast<uint32_t>(h.size()), &b_), "ckmpm_batch_create",
                      nullptr);
    }
    ~Batch() {
        if (b_) ckmpm_batch_destroy(b_);
    }
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;
    void Run(uint32_t substeps, Mode mode = kFused) {
        detail::Check(ckmpm_batch_run(b_, substeps, mode), "ckmpm_batch_run", ckmpm_batch_last_error(b_));
    }
    void Synchronize() {
        detail::Check(ckmpm_batch_synchronize(b_), "ckmpm_batch_synchronize", ckmpm_batch_last_error(b_));
    }

   private:
    ckmpm_batch* b_ = nullptr;
};

}  // namespace ckmpm

#endif /* CKMPM_HPP */
