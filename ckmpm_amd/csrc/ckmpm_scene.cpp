// ckmpm_scene: a scene main over the C++ facade (include/ckmpm.hpp).
//
// The reference runs each scene from a compiled main that builds its models (positions
// read from raw float32 xyz assets, e.g. tests/chimney_ball/test_chimney_ball.cuh:140-188)
// and calls Initialize, InitializeParticle, InitialSetup and Simulate
// (tests/colliding_sphere/test_colliding_sphere.cu:5-19).  This program does the same for
// a scene given on the command line, so the reference's assets run without a recompile.
// Two colliding spheres, on one command line:
//
//   ckmpm_scene --domain 16,16,16 --dx 0.015625 --dt 1e-4
//       --material fc --E 1e4 --nu 0.3 --mass 2.4e-4 --volume 4.8e-7 --velocity 1,0,0
//       --positions left.bin
//       --material fc --E 1e4 --nu 0.3 --mass 2.4e-4 --volume 4.8e-7 --velocity -1,0,0
//       --positions right.bin
//       --substeps 100 --out result/model
//
// Every --positions closes one model with the material options given before it.  The
// final positions of model i go to <out>i.bin in the same raw format (storage order).
// With --conservation DIR the per-substep conservation metric of model i goes to
// DIR/conservationMetric_model_<i>.bin, the reference's file (mpm_engine.cuh:494-603).
// With --export DIR and --frames, the positions of model i after frame f go to
// DIR/model_<i>_particle_frame_<f>.bgeo (Houdini binary geometry, as the reference's
// partio export, mpm_engine.cuh:446-452) and DIR/model_<i>_particle_frame_<f>.txt, raw
// float32 xyz, as its exportParticlePositionAsBinary files (mpm_engine.cuh:432-444,
// written at the end of every frame, :880-883).
// One JSON line on stdout reports the run.  Exit status 0 on success, 2 on a usage error,
// 1 on a library error (its CKMPM_ERR_* code and message on stderr).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "ckmpm.hpp"

namespace {

const char kUsage[] =
    "usage: ckmpm_scene --domain DX,DY,DZ --dx H [scene options] MODEL... (--substeps K | --frames F)\n"
    "scene options:\n"
    "  --offset OX,OY,OZ   domain offset in blocks (0,0,0)\n"
    "  --dt T              default and maximum dt (1e-4)\n"
    "  --gravity G         acceleration along y (0)\n"
    "  --mass-clamp M      grid nodes with mass <= M get zero velocity (0)\n"
    "  --cfl C             CFL number of the frame loop's dt rule (0.5)\n"
    "  --fps F             frames per second of the frame loop (48)\n"
    "  --capacity B        block capacity of each layout (65536)\n"
    "  --autotune K        first pick each model's G2P2G workgroup size (advances the scene)\n"
    "  --out PREFIX        write PREFIX<i>.bin: final positions of model i, float32 xyz\n"
    "  --conservation DIR  write DIR/conservationMetric_model_<i>.bin: 13 floats per substep\n"
    "  --walls sticky|slip|none   grid walls within 8 cells of the faces (sticky)\n"
    "  --region fixed,X0,Y0,Z0,X1,Y1,Z1 | velocity,<box>,VX,VY,VZ | rotation,<box>,WX,WY,WZ,CX,CY,CZ\n"
    "                      grid region on cells [X0,X1) x [Y0,Y1) x [Z0,Z1), applied in order\n"
    "                      before the walls; any --walls / --region replaces the default rule\n"
    "  --export DIR        with --frames: after frame f write DIR/model_<i>_particle_frame_<f>.bgeo\n"
    "                      (Houdini binary geometry) and .txt (raw float32 xyz), as the reference\n"
    "a model (options before the --positions that closes it):\n"
    "  --material fc|nh|dp|fluid|vm|nacc\n"
    "  --lambda L --mu M   or  --E E --nu NU   (solids)\n"
    "  --friction DEG      Drucker-Prager friction angle (30)\n"
    "  --bulk K --gamma G --viscosity V --J0 J   (fluid)\n"
    "  --yield S           von Mises yield stress\n"
    "  --alpha0 A --beta B --xi X --M M --hardening 0|1   (Cam-clay; hardening 1)\n"
    "  --mass M --volume V particle mass and volume\n"
    "  --velocity VX,VY,VZ one initial velocity for all particles (0,0,0)\n"
    "  --velocities FILE   per-particle initial velocities, float32 xyz\n"
    "  --positions FILE    particle positions, float32 xyz (the reference's .bin assets)\n";

struct UsageError {
    std::string what;
};

double Number(const char* opt, const char* s) {
    char* end = nullptr;
    const double v = std::strtod(s, &end);
    if (end == s || *end != '\0') throw UsageError{std::string(opt) + " needs a number, got '" + s + "'"};
    return v;
}

std::vector<double> Numbers(const char* opt, const char* s, size_t count) {
    std::vector<double> out;
    std::string cur;
    for (const char* p = s;; ++p) {
        if (*p == ',' || *p == '\0') {
            out.push_back(Number(opt, cur.c_str()));
            cur.clear();
            if (*p == '\0') break;
        } else {
            cur += *p;
        }
    }
    if (out.size() != count) throw UsageError{std::string(opt) + " needs " + std::to_string(count) + " comma-separated numbers"};
    return out;
}

struct PendingModel {
    std::string material;
    double lambda = -1, mu = -1, E = -1, nu = -1, friction = 30;
    double bulk = 0, gamma = 0, viscosity = 0, J0 = 1;
    double yield = 0, alpha0 = 0, beta = 0, xi = 0, M = 0, hardening = 1;
    double mass = 0, volume = 0;
    std::vector<float> velocities{0.0f, 0.0f, 0.0f};
};

ckmpm::Model Close(const PendingModel& p, const std::string& positions) {
    ckmpm::Model m;
    if (p.material != "fc" && p.material != "nh" && p.material != "dp" && p.material != "fluid" && p.material != "vm" &&
        p.material != "nacc")
        throw UsageError{"--positions " + positions + ": --material must be fc, nh, dp, fluid, vm or nacc"};
    float lambda = static_cast<float>(p.lambda), mu = static_cast<float>(p.mu);
    if (p.material != "fluid") {
        if (p.E > 0) {
            const auto lm = ckmpm::LameParameters(p.E, p.nu);
            lambda = lm[0];
            mu = lm[1];
        } else if (p.lambda < 0 || p.mu < 0) {
            throw UsageError{"a solid needs --lambda and --mu, or --E and --nu"};
        }
    }
    if (p.material == "fc") m.material = ckmpm::Material::FixedCorotated(lambda, mu);
    else if (p.material == "nh") m.material = ckmpm::Material::NeoHookean(lambda, mu);
    else if (p.material == "dp") m.material = ckmpm::Material::DruckerPrager(lambda, mu, static_cast<float>(p.friction));
    else if (p.material == "vm") m.material = ckmpm::Material::VonMises(lambda, mu, static_cast<float>(p.yield));
    else if (p.material == "nacc")
        m.material = ckmpm::Material::CamClay(lambda, mu, static_cast<float>(p.alpha0), static_cast<float>(p.beta),
                                              static_cast<float>(p.xi), static_cast<float>(p.M), p.hardening != 0);
    else
        m.material = ckmpm::Material::Fluid(static_cast<float>(p.bulk), static_cast<float>(p.gamma),
                                            static_cast<float>(p.viscosity), static_cast<float>(p.J0));
    if (!(p.mass > 0) || !(p.volume > 0)) throw UsageError{"--positions " + positions + ": --mass and --volume must be > 0"};
    m.mass = static_cast<float>(p.mass);
    m.volume = static_cast<float>(p.volume);
    m.positions = ckmpm::ReadRawPositions(positions);
    m.velocities = p.velocities;
    return m;
}

int Run(int argc, char** argv) {
    ckmpm::Domain domain;
    ckmpm::Config config;
    bool have_domain = false;
    long substeps = -1, frames = -1, autotune = 0;
    std::string out, cons_dir, export_dir;
    std::vector<ckmpm::Model> models;
    PendingModel pend;
    ckmpm::Boundary boundary;
    bool custom_boundary = false;
    for (int i = 1; i < argc; ++i) {
        const char* opt = argv[i];
        if (!std::strcmp(opt, "--help") || !std::strcmp(opt, "-h")) {
            std::fputs(kUsage, stdout);
            return 0;
        }
        if (i + 1 >= argc) throw UsageError{std::string(opt) + " needs a value"};
        const char* val = argv[++i];
        const std::string o(opt);
        if (o == "--domain" || o == "--offset") {
            const auto v = Numbers(opt, val, 3);
            for (int k = 0; k < 3; ++k) (o == "--domain" ? domain.dim : domain.offset)[k] = static_cast<int32_t>(v[k]);
            have_domain |= o == "--domain";
        } else if (o == "--dx") config.dx = static_cast<float>(Number(opt, val));
        else if (o == "--dt") config.dt = static_cast<float>(Number(opt, val));
        else if (o == "--gravity") config.gravity = static_cast<float>(Number(opt, val));
        else if (o == "--mass-clamp") config.mass_clamp = static_cast<float>(Number(opt, val));
        else if (o == "--cfl") config.cfl = static_cast<float>(Number(opt, val));
        else if (o == "--fps") config.fps = static_cast<uint32_t>(Number(opt, val));
        else if (o == "--capacity") config.max_active_block_count = static_cast<uint32_t>(Number(opt, val));
        else if (o == "--substeps") substeps = static_cast<long>(Number(opt, val));
        else if (o == "--frames") frames = static_cast<long>(Number(opt, val));
        else if (o == "--autotune") autotune = static_cast<long>(Number(opt, val));
        else if (o == "--out") out = val;
        else if (o == "--conservation") cons_dir = val;
        else if (o == "--export") export_dir = val;
        else if (o == "--walls") {
            const std::string w(val);
            if (w == "sticky") boundary.d.walls = CKMPM_WALLS_STICKY;
            else if (w == "slip") boundary.d.walls = CKMPM_WALLS_SLIP;
            else if (w == "none") boundary.d.walls = CKMPM_WALLS_NONE;
            else throw UsageError{"--walls must be sticky, slip or none"};
            custom_boundary = true;
        } else if (o == "--region") {
            const std::string spec(val);
            const size_t comma = spec.find(',');
            const std::string kind = spec.substr(0, comma);
            const std::string rest = comma == std::string::npos ? "" : spec.substr(comma + 1);
            const size_t want = kind == "fixed" ? 6 : kind == "velocity" ? 9 : kind == "rotation" ? 12 : 0;
            if (!want) throw UsageError{"--region kind must be fixed, velocity or rotation"};
            const auto v = Numbers(opt, rest.c_str(), want);
            const std::array<int32_t, 3> lo{static_cast<int32_t>(v[0]), static_cast<int32_t>(v[1]), static_cast<int32_t>(v[2])};
            const std::array<int32_t, 3> hi{static_cast<int32_t>(v[3]), static_cast<int32_t>(v[4]), static_cast<int32_t>(v[5])};
            auto f3 = [&](size_t i) {
                return std::array<float, 3>{static_cast<float>(v[i]), static_cast<float>(v[i + 1]), static_cast<float>(v[i + 2])};
            };
            if (kind == "fixed") boundary.Fixed(lo, hi);
            else if (kind == "velocity") boundary.Velocity(lo, hi, f3(6));
            else boundary.Rotation(lo, hi, f3(6), f3(9));
            custom_boundary = true;
        }
        else if (o == "--material") pend.material = val;
        else if (o == "--lambda") pend.lambda = Number(opt, val);
        else if (o == "--mu") pend.mu = Number(opt, val);
        else if (o == "--E") pend.E = Number(opt, val);
        else if (o == "--nu") pend.nu = Number(opt, val);
        else if (o == "--friction") pend.friction = Number(opt, val);
        else if (o == "--bulk") pend.bulk = Number(opt, val);
        else if (o == "--gamma") pend.gamma = Number(opt, val);
        else if (o == "--viscosity") pend.viscosity = Number(opt, val);
        else if (o == "--J0") pend.J0 = Number(opt, val);
        else if (o == "--yield") pend.yield = Number(opt, val);
        else if (o == "--alpha0") pend.alpha0 = Number(opt, val);
        else if (o == "--beta") pend.beta = Number(opt, val);
        else if (o == "--xi") pend.xi = Number(opt, val);
        else if (o == "--M") pend.M = Number(opt, val);
        else if (o == "--hardening") pend.hardening = Number(opt, val);
        else if (o == "--mass") pend.mass = Number(opt, val);
        else if (o == "--volume") pend.volume = Number(opt, val);
        else if (o == "--velocity") {
            const auto v = Numbers(opt, val, 3);
            pend.velocities.assign({static_cast<float>(v[0]), static_cast<float>(v[1]), static_cast<float>(v[2])});
        } else if (o == "--velocities") pend.velocities = ckmpm::ReadRawPositions(val);
        else if (o == "--positions") {
            models.push_back(Close(pend, val));
            pend = PendingModel{};
        } else throw UsageError{"unknown option " + o};
    }
    if (!have_domain || !(config.dx > 0)) throw UsageError{"--domain and --dx are required"};
    if (models.empty()) throw UsageError{"no model: give --positions FILE after each model's options"};
    if ((substeps < 0) == (frames < 0)) throw UsageError{"give exactly one of --substeps and --frames"};
    if (!export_dir.empty() && frames < 0) throw UsageError{"--export writes frames: use it with --frames"};

    ckmpm::Engine engine(domain, config);
    engine.InitializeParticle(models);
    if (custom_boundary) engine.SetBoundary(boundary);
    engine.InitialSetup();
    if (autotune > 0) engine.Autotune(static_cast<uint32_t>(autotune));
    engine.Synchronize();
    uint64_t particles = 0;
    for (const ckmpm::Model& m : models) particles += m.count();
    // conservation records: appended to the files after the substeps, or after every frame
    // (the frame loop's substep count depends on the CFL dt)
    struct Closer {
        void operator()(std::FILE* f) const { std::fclose(f); }
    };
    std::vector<std::unique_ptr<std::FILE, Closer>> cons;
    for (size_t i = 0; !cons_dir.empty() && i < models.size(); ++i) {
        const std::string path = cons_dir + "/conservationMetric_model_" + std::to_string(i) + ".bin";
        cons.emplace_back(std::fopen(path.c_str(), "wb"));
        if (!cons.back()) throw ckmpm::Error(CKMPM_ERR_INVALID, "cannot create " + path);
    }
    const uint32_t cons_cap = substeps >= 0 ? static_cast<uint32_t>(std::max(substeps, 1L)) : 1u << 20;
    auto flush_cons = [&]() {
        for (size_t i = 0; i < cons.size(); ++i) {
            const std::vector<float> r = engine.Conservation(static_cast<int32_t>(i));
            if (!r.empty() && std::fwrite(r.data(), 4, r.size(), cons[i].get()) != r.size())
                throw ckmpm::Error(CKMPM_ERR_INVALID, "cannot write the conservation records");
        }
        engine.SetConservation(cons_cap);  // clears the log
    };
    if (!cons.empty()) engine.SetConservation(cons_cap);
    const auto t0 = std::chrono::steady_clock::now();
    const uint64_t before = engine.Counters().substeps;
    if (substeps >= 0) {
        engine.Substeps(static_cast<uint32_t>(substeps));
    } else {
        for (long f = 0; f < frames; ++f) {
            engine.Simulate(1);
            if (!cons.empty()) flush_cons();
            for (size_t i = 0; !export_dir.empty() && i < models.size(); ++i) {
                const std::string stem =
                    export_dir + "/model_" + std::to_string(i) + "_particle_frame_" + std::to_string(f);
                const std::vector<float> xyz = engine.Positions(static_cast<int32_t>(i));
                ckmpm::WriteRawPositions(stem + ".txt", xyz);
                ckmpm::WriteBgeoPositions(stem + ".bgeo", xyz);
            }
        }
    }
    engine.Synchronize();
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (substeps >= 0 && !cons.empty()) flush_cons();
    for (auto& f : cons)
        if (std::fclose(f.release()) != 0) throw ckmpm::Error(CKMPM_ERR_INVALID, "cannot write the conservation records");
    const ckmpm_counters c = engine.Counters();
    const uint64_t ran = c.substeps - before;
    if (!out.empty())
        for (size_t i = 0; i < models.size(); ++i)
            ckmpm::WriteRawPositions(out + std::to_string(i) + ".bin", engine.Positions(static_cast<int32_t>(i)));
    std::printf(
        "{\"particles\": %llu, \"models\": %zu, \"substeps\": %llu, \"seconds\": %.6f, "
        "\"particle_substeps_per_s\": %.6e, \"blocks\": [%u, %u, %u], \"max_speed\": %.6e, \"flags\": %u}\n",
        static_cast<unsigned long long>(particles), models.size(), static_cast<unsigned long long>(ran), sec,
        sec > 0 ? static_cast<double>(particles) * static_cast<double>(ran) / sec : 0.0, c.active_blocks,
        c.neighbor_blocks, c.exterior_blocks, static_cast<double>(c.max_speed), c.flags);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        return Run(argc, argv);
    } catch (const UsageError& e) {
        std::fprintf(stderr, "ckmpm_scene: %s\n%s", e.what.c_str(), kUsage);
        return 2;
    } catch (const ckmpm::Error& e) {
        std::fprintf(stderr, "ckmpm_scene: %s (CKMPM_ERR %d)\n", e.what(), e.code());
        return 1;
    }
}
