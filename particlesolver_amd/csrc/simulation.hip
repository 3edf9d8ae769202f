// simulation.hip — the C++ Simulation facade (the reference CPU app's
// Simulation surface) over the C-ABI.  The scene builders restate the
// reference's scene functions with SplitMix64 in place of frand(); the
// lattices, seeds and arithmetic are those of particlesolver_amd/scenes.py
// (reference_two_fluids, reference_gas, reference_smoke,
// reference_volcano), so a scene built here equals the one built there.

#include "simulation.hpp"

#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

namespace ps {

namespace {

// the j-th of `count` FP64 uniforms in [0, 1) with 24 random bits:
// SplitMix64 of seed + (j + 1) * golden, top 24 bits
double uniform24(uint64_t seed, uint64_t j)
{
    uint64_t z = seed + (j + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 40) / 16777216.0;
}

// the CPU app's scene lattice: x from x0 while x < x1, y from y0 while
// y < y1 (x outer), step delta, each point plus .2 (u - .5) per axis
std::vector<Float2> lattice(double x0, double x1, double y0, double y1, double delta, uint64_t seed)
{
    std::vector<double> pts;
    for (double x = x0; x < x1; x += delta)
        for (double y = y0; y < y1; y += delta) {
            pts.push_back(x);
            pts.push_back(y);
        }
    std::vector<Float2> out(pts.size() / 2);
    for (size_t k = 0; k < out.size(); k++) {
        const double ux = uniform24(seed, 2 * k), uy = uniform24(seed, 2 * k + 1);
        out[k] = {(float)(pts[2 * k] + 0.2 * (ux - 0.5)), (float)(pts[2 * k + 1] + 0.2 * (uy - 0.5))};
    }
    return out;
}

std::vector<float> xyz_of(const std::vector<Float2> &pos)
{
    std::vector<float> xyz(3 * pos.size());
    for (size_t k = 0; k < pos.size(); k++) {
        xyz[3 * k] = pos[k].x;
        xyz[3 * k + 1] = pos[k].y;
        xyz[3 * k + 2] = 0.0f;
    }
    return xyz;
}

// Particle::init (cpu/src/particle.h:37-48): imass = 1/mass, or -mass for
// mass <= 0
float inverse_mass(double mass) { return (float)(mass <= 0.0 ? -mass : 1.0 / mass); }

}  // namespace

Simulation::Simulation(unsigned maxParticles) : m_maxParticles(maxParticles) {}

Simulation::~Simulation() { clear(); }

void Simulation::check(int rc) const
{
    if (rc != PS_OK) throw Error(rc, ps_last_error(m_sys));
}

void Simulation::clear()
{
    if (m_sys) ps_destroy(m_sys);
    m_sys = nullptr;
    m_objects = 0;
    m_imass.clear();
}

ps_params Simulation::params() const
{
    // the reference's constants (particle.h:6-7, totalfluidconstraint.h:22-30,
    // simulation.h:11, boundaryconstraint.cpp:19: jitter frand() * .003)
    ps_params p;
    std::memset(&p, 0, sizeof(p));
    p.abi_version = PS_ABI_VERSION;
    p.dims = 2;
    p.radius = 0.25f;
    p.h = 2.0f;
    p.cell = 2.0f;
    p.collide_dist = 2.0f;
    p.gravity[0] = m_gravity.x;
    p.gravity[1] = m_gravity.y;
    p.box_min[0] = m_xBoundaries.x;
    p.box_max[0] = m_xBoundaries.y;
    p.box_min[1] = m_yBoundaries.x;
    p.box_max[1] = m_yBoundaries.y;
    p.box_min[2] = -1.0f;
    p.box_max[2] = 1.0f;
    for (int a = 0; a < 3; a++) {
        p.origin[a] = p.box_min[a];
        p.grid[a] = 1;
    }
    p.iterations = 3;
    p.relaxation = 0.01f;
    p.k_corr = 0.1f;
    p.n_corr = 4;
    p.dq_corr = 0.2f;
    p.wall_jitter = (float)(0.003 / 0.25);
    p.shape_stiffness = 1.0f;
    p.max_particles = m_maxParticles;
    p.max_neighbors = 4;
    p.seed = 0x5EED0003ull;
    p.flags = PS_FLAG_CPU_REFERENCE | PS_FLAG_SLEEPING;
    return p;
}

ps_system *Simulation::system()
{
    if (!m_sys) {
        const ps_params p = params();
        check(ps_create(&p, &m_sys));
    }
    return m_sys;
}

ps_system *Simulation::handle() { return system(); }

void Simulation::setBoundaries(Float2 x, Float2 y)
{
    m_xBoundaries = x;
    m_yBoundaries = y;
    if (m_sys) {
        const ps_params p = params();
        check(ps_set_params(m_sys, &p));
    }
}

void Simulation::setGravity(Float2 g)
{
    m_gravity = g;
    if (m_sys) {
        const ps_params p = params();
        check(ps_set_params(m_sys, &p));
    }
}

void Simulation::addStaticParticles(const std::vector<Float2> &pos)
{
    const std::vector<float> xyz = xyz_of(pos), zero(pos.size(), 0.0f);
    const std::vector<int32_t> phase(pos.size(), PS_PHASE_SOLID);
    check(ps_add_particles(system(), (uint32_t)pos.size(), xyz.data(), nullptr, zero.data(),
                           phase.data(), zero.data(), nullptr));
    m_imass.insert(m_imass.end(), pos.size(), 0.0f);
}

int Simulation::createFluid(const std::vector<Float2> &pos, double density, double mass)
{
    const std::vector<float> xyz = xyz_of(pos), w(pos.size(), inverse_mass(mass));
    check(ps_add_fluid_object(system(), (uint32_t)pos.size(), xyz.data(), nullptr, w.data(),
                              (float)density, nullptr));
    m_imass.insert(m_imass.end(), w.begin(), w.end());
    return m_objects++;
}

int Simulation::createGas(const std::vector<Float2> &pos, double density, bool open, double mass)
{
    const std::vector<float> xyz = xyz_of(pos), w(pos.size(), inverse_mass(mass));
    check(ps_add_gas_object(system(), (uint32_t)pos.size(), xyz.data(), nullptr, w.data(),
                            (float)density, open ? 1 : 0, nullptr));
    m_imass.insert(m_imass.end(), w.begin(), w.end());
    return m_objects++;
}

unsigned Simulation::createSmokeEmitter(Float2 posn, double particlesPerSec, int gasObject,
                                        unsigned maxMarkers)
{
    const float p[3] = {posn.x, posn.y, 0.0f};
    uint32_t e = 0;
    check(ps_add_smoke_emitter(system(), p, particlesPerSec, gasObject, maxMarkers, &e));
    return e;
}

unsigned Simulation::createFluidEmitter(Float2 posn, double particlesPerSec, int fluidObject)
{
    const float p[3] = {posn.x, posn.y, 0.0f};
    uint32_t e = 0;
    check(ps_add_fluid_emitter(system(), p, particlesPerSec, fluidObject, &e));
    return e;
}

void Simulation::init(SimulationType type)
{
    switch (type) {
    case FLUID_TEST: initFluid(); break;
    case SMOKE_OPEN_TEST: initSmokeOpen(); break;
    case SMOKE_CLOSED_TEST: initSmokeClosed(); break;
    case VOLCANO_TEST: initVolcano(); break;
    default:
        throw Error(PS_ERR_UNSUPPORTED,
                    "scene " + std::to_string((int)type) +
                        " needs rigid bodies, SDF contacts or 2D distance constraints");
    }
}

void Simulation::initFluid()
{
    clear();
    const double scale = 4.0, delta = 0.7;
    setGravity({0.0f, -9.8f});
    setBoundaries({(float)(-2 * scale), (float)(2 * scale)}, {(float)(-2 * scale), (float)(10 * scale)});
    for (int d = 0; d < 2; d++) {
        const double start = -2.0 * scale + 4.0 * scale * (d / 2.0);
        createFluid(lattice(start, start + 4.0 * scale / 2.0, -2.0 * scale, scale, delta, 0x5EED0005ull + d),
                    1.0 + 0.75 * d);
    }
}

void Simulation::initGas()
{
    clear();
    const double delta = 0.7;
    setGravity({0.0f, -9.8f});
    setBoundaries({-4.0f, 4.0f}, {-4.0f, 20.0f});
    double scale = 2.0;
    for (int d = 0; d < 2; d++) {
        const double start = -2.0 * scale + 4.0 * scale * (d / 2.0);
        createGas(lattice(start, start + 4.0 * scale / 2.0, -2.0 * scale, 2.0 * scale, delta,
                          0x5EED0008ull + d),
                  0.75 + 3.0 * d, false);
    }
    scale = 3.0;
    for (int d = 0; d < 2; d++) {
        const double start = -2.0 * scale + 4.0 * scale * (d / 2.0);
        createFluid(lattice(start, start + 4.0 * scale / 2.0, -2.0 * scale + 10.0, 2.0 * scale + 10.0,
                            delta, 0x5EED0008ull + 2 + d),
                    4.0 + 0.75 * (d + 1));
    }
}

void Simulation::initSmokeOpen()
{
    clear();
    const double scale = 2.0, delta = 0.63;
    setGravity({0.0f, -9.8f});
    setBoundaries({(float)(-3 * scale), (float)(3 * scale)}, {(float)(-2 * scale), (float)(100 * scale)});
    const int gs = createGas(lattice(-2.0 * scale, 2.0 * scale, -2.0 * scale, 2.0 * scale, delta,
                                     0x5EED0009ull),
                             1.5, true);
    createSmokeEmitter({0.0f, (float)(-2 * scale + 1)}, 15.0, gs);
}

void Simulation::initSmokeClosed()
{
    clear();
    const double scale = 2.0, delta = 0.63;
    setGravity({0.0f, -9.8f});
    setBoundaries({(float)(-2 * scale), (float)(2 * scale)}, {(float)(-2 * scale), (float)(2 * scale)});
    createGas(lattice(-2.0 * scale, 2.0 * scale, -2.0 * scale, 2.0 * scale, delta, 0x5EED0009ull), 1.5,
              false);
    createSmokeEmitter({0.0f, (float)(-2 * scale + 1)}, 15.0, -1);
}

void Simulation::initVolcano()
{
    clear();
    const double scale = 10.0;
    setGravity({0.0f, -9.8f});
    setBoundaries({(float)(-2 * scale), (float)(2 * scale)}, {0.0f, (float)(10 * scale)});
    std::vector<Float2> walls;
    for (double x = 1.0; x <= scale; x += 0.2) {
        walls.push_back({(float)-x, (float)(scale - x)});
        walls.push_back({(float)x, (float)(scale - x)});
    }
    addStaticParticles(walls);
    std::vector<double> pts;
    for (double y = 0.0; y < scale - 1.0; y += 0.8)
        for (double x = 0.0; x < scale - y - 1.0; x += 0.8) {
            pts.insert(pts.end(), {x, y, -x, y});
        }
    std::vector<Float2> fluid(pts.size() / 2);
    for (size_t k = 0; k < fluid.size(); k++) {
        const double ux = uniform24(0x5EED000Aull, 2 * k), uy = uniform24(0x5EED000Aull, 2 * k + 1);
        fluid[k] = {(float)(pts[2 * k] + 0.2 * (ux - 0.5)), (float)(pts[2 * k + 1] + 0.2 * (uy - 0.5))};
    }
    const int fs = createFluid(fluid, 1.0, 1.1);
    createFluidEmitter({0.0f, 0.0f}, scale * 4, fs);
}

void Simulation::tick(double seconds) { check(ps_step(system(), (float)seconds, 1)); }

int Simulation::getNumParticles() { return m_sys ? (int)ps_num_particles(m_sys) : 0; }

double Simulation::getKineticEnergy()
{
    const int n = getNumParticles();
    if (!n) return 0.0;
    std::vector<float> v(3 * (size_t)n);
    check(ps_read_velocities(m_sys, 0, (uint32_t)n, v.data()));
    // phases tell the particles a fluid emitter turned solid (imass 0)
    std::vector<int32_t> phase(n);
    int32_t *dph = nullptr;
    if (hipMalloc((void **)&dph, sizeof(int32_t) * n) != hipSuccess)
        throw Error(PS_ERR_OUT_OF_MEMORY, "getKineticEnergy: device buffer");
    int rc = ps_export_device(m_sys, PS_FIELD_PHASES, 0, (uint32_t)n, dph);
    if (rc == PS_OK) rc = ps_synchronize(m_sys);
    if (rc == PS_OK && hipMemcpy(phase.data(), dph, sizeof(int32_t) * n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = PS_ERR_DEVICE;
    (void)hipFree(dph);
    check(rc);
    double e = 0.0;
    for (int i = 0; i < n; i++) {
        // particles added here carry their own inverse mass; emitted ones 1
        const double w = (size_t)i < m_imass.size() ? m_imass[i] : 1.0;
        if (phase[i] == PS_PHASE_SOLID || w == 0.0) continue;
        const double vx = v[3 * (size_t)i], vy = v[3 * (size_t)i + 1];
        e += (0.5 * (vx * vx + vy * vy)) / w;
    }
    return e;
}

void Simulation::readPositions(std::vector<float> &xyz)
{
    const int n = getNumParticles();
    xyz.resize(3 * (size_t)n);
    if (n) check(ps_read_positions(m_sys, 0, (uint32_t)n, xyz.data()));
}

void Simulation::readVelocities(std::vector<float> &xyz)
{
    const int n = getNumParticles();
    xyz.resize(3 * (size_t)n);
    if (n) check(ps_read_velocities(m_sys, 0, (uint32_t)n, xyz.data()));
}

void Simulation::readMarkers(unsigned emitter, std::vector<float> &xyz)
{
    uint32_t c = 0;
    check(ps_read_markers(system(), emitter, &c, nullptr));
    xyz.resize(3 * (size_t)c);
    if (c) check(ps_read_markers(m_sys, emitter, &c, xyz.data()));
}

}  // namespace ps
