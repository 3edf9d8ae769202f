// particle_system.hip — C++ ParticleSystem facade over the C-ABI.
// Scene builders follow gpu/src/particlesystem.cpp:357-698; the lattice
// arithmetic (counts from (int)ceil(extent)/spacing, jitter r*0.01) is the
// reference's, so a scene built here has the reference's particle layout.

#include "particle_system.hpp"

#include <algorithm>
#include <cmath>

namespace ps {

namespace {

ps_params facade_params(float r, unsigned maxParticles, Int3 lo, Int3 hi, int iterations,
                        float h)
{
    ps_params p;
    ps_params_default(&p);
    p.radius = r;
    p.h = h;
    p.collide_dist = r * 2.001f;  // integration_kernel.cuh:319
    p.cell = std::max(h, p.collide_dist);
    const int lov[3] = {lo.x, lo.y, lo.z}, hiv[3] = {hi.x, hi.y, hi.z};
    for (int a = 0; a < 3; a++) {
        p.box_min[a] = (float)lov[a];
        p.box_max[a] = (float)hiv[a];
        p.origin[a] = (float)lov[a];
        const double ext = (double)hiv[a] - (double)lov[a];
        p.grid[a] = (uint32_t)std::max(1.0, std::ceil(ext / (double)p.cell));
    }
    p.iterations = (uint32_t)std::max(1, iterations);
    p.max_particles = maxParticles;
    p.max_neighbors = 500;  // MAX_FLUID_NEIGHBORS, integration_kernel.cuh:23
    return p;
}

}  // namespace

ParticleSystem::ParticleSystem(float particleRadius, UInt3 /*gridSize*/, unsigned maxParticles,
                               Int3 minBounds, Int3 maxBounds, int iterations, float h)
    : ParticleSystem(facade_params(particleRadius, maxParticles, minBounds, maxBounds,
                                   iterations, h))
{
    m_minBounds = minBounds;
    m_maxBounds = maxBounds;
}

ParticleSystem::ParticleSystem(const ps_params &params) : m_params(params)
{
    m_minBounds = {(int)params.box_min[0], (int)params.box_min[1], (int)params.box_min[2]};
    m_maxBounds = {(int)params.box_max[0], (int)params.box_max[1], (int)params.box_max[2]};
    m_rng = params.seed;
    check(ps_create(&m_params, &m_sys));
}

ParticleSystem::~ParticleSystem() { ps_destroy(m_sys); }

void ParticleSystem::check(int rc) const
{
    if (rc != PS_OK) throw Error(rc, ps_last_error(m_sys));
}

float ParticleSystem::frand()
{
    // SplitMix64, top 24 bits -> [0, 1)
    uint64_t z = (m_rng += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (float)((double)(uint32_t)(z >> 40) / 16777216.0);
}

unsigned ParticleSystem::getNumParticles() const { return ps_num_particles(m_sys); }

void ParticleSystem::update(float deltaTime)
{
    deltaTime = std::min(deltaTime, 0.05f);
    if (getNumParticles() > 0 && deltaTime > 0.0f) check(ps_step(m_sys, deltaTime, 1));
    // addNewStuff (particlesystem.cpp:249-253): queued particles, then fluids
    while (m_particlesToAdd.size() >= 7) {
        std::vector<float> pos(3), vel(3), w(1), ro(1, 1.5f);
        for (int a = 0; a < 3; a++) { pos[a] = m_particlesToAdd.front(); m_particlesToAdd.pop_front(); }
        for (int a = 0; a < 3; a++) { vel[a] = m_particlesToAdd.front(); m_particlesToAdd.pop_front(); }
        w[0] = 1.0f / m_particlesToAdd.front();
        m_particlesToAdd.pop_front();
        if (getNumParticles() == m_params.max_particles) continue;  // particlesystem.cpp:303-304
        addParticleMultiple(pos, vel, w, ro, std::vector<int32_t>(1, PS_PHASE_SOLID), true);
    }
    if (m_fluidsToAdd.size() >= 5) {
        std::vector<float> pos, vel, w, ro;
        std::vector<int32_t> ph;
        while (m_fluidsToAdd.size() >= 5) {
            for (int a = 0; a < 3; a++) { pos.push_back(m_fluidsToAdd.front()); m_fluidsToAdd.pop_front(); }
            vel.insert(vel.end(), {0.0f, -1.0f, 0.0f});  // particlesystem.cpp:326
            w.push_back(1.0f / m_fluidsToAdd.front());
            m_fluidsToAdd.pop_front();
            ro.push_back(m_fluidsToAdd.front());
            m_fluidsToAdd.pop_front();
            ph.push_back(PS_PHASE_FLUID);
        }
        addParticleMultiple(pos, vel, w, ro, ph);
    }
}

// particlesystem.cpp:333-347: a block that would reach maxParticles is
// dropped whole, without an error, as there (false then; the builders then
// add none of its constraints)
bool ParticleSystem::addParticleMultiple(const std::vector<float> &pos,
                                         const std::vector<float> &vel,
                                         const std::vector<float> &w,
                                         const std::vector<float> &ro,
                                         const std::vector<int32_t> &phase, bool single)
{
    const uint32_t n = (uint32_t)w.size();
    if (n == 0) return true;
    if (!single && (uint64_t)getNumParticles() + n >= m_params.max_particles) return false;
    check(ps_add_particles(m_sys, n, pos.data(), vel.data(), w.data(), phase.data(), ro.data(),
                           nullptr));
    return true;
}

// particlesystem.cpp:357-405: lattice spacing 2.5 r, jitter r * 0.01
void ParticleSystem::addFluid(Int3 ll, Int3 ur, float mass, float density, Float3 /*color*/)
{
    const float r = m_params.radius;
    const float jitter = r * 0.01f, distance = r * 2.5f;
    const int cx = (int)((float)(int)std::ceil((float)(ur.x - ll.x)) / distance);
    const int cy = (int)((float)(int)std::ceil((float)(ur.y - ll.y)) / distance);
    const int cz = m_params.dims == 2 ? 1 : (int)((float)(int)std::ceil((float)(ur.z - ll.z)) / distance);
    std::vector<float> pos, vel, w, ro;
    std::vector<int32_t> ph;
    for (int z = 0; z < cz; z++)
        for (int y = 0; y < cy; y++)
            for (int x = 0; x < cx; x++) {
                const float jx = (frand() * 2.0f - 1.0f) * jitter;
                const float jy = (frand() * 2.0f - 1.0f) * jitter;
                const float jz = (frand() * 2.0f - 1.0f) * jitter;
                pos.push_back(ll.x + x * distance + jx);
                pos.push_back(ll.y + y * distance + jy);
                pos.push_back(m_params.dims == 2 ? m_params.box_min[2] : ll.z + z * distance + jz);
                vel.insert(vel.end(), {0.0f, 0.0f, 0.0f});
                w.push_back(1.0f / mass);
                ro.push_back(density);
                ph.push_back(PS_PHASE_FLUID);
            }
    addParticleMultiple(pos, vel, w, ro, ph);
}

// particlesystem.cpp:408-458: spacing 2.002 r, optional jitter, SOLID phase
void ParticleSystem::addParticleGrid(Int3 ll, Int3 ur, float mass, bool addJitter)
{
    const float r = m_params.radius;
    const float jitter = addJitter ? r * 0.01f : 0.0f, distance = r * 2.002f;
    const int cx = (int)((float)(int)std::ceil((float)(ur.x - ll.x)) / distance);
    const int cy = (int)((float)(int)std::ceil((float)(ur.y - ll.y)) / distance);
    const int cz = m_params.dims == 2 ? 1 : (int)((float)(int)std::ceil((float)(ur.z - ll.z)) / distance);
    std::vector<float> pos, vel, w, ro;
    std::vector<int32_t> ph;
    for (int z = 0; z < cz; z++)
        for (int y = 0; y < cy; y++)
            for (int x = 0; x < cx; x++) {
                const float jx = (frand() * 2.0f - 1.0f) * jitter;
                const float jy = (frand() * 2.0f - 1.0f) * jitter;
                const float jz = (frand() * 2.0f - 1.0f) * jitter;
                pos.push_back(ll.x + x * distance + jx);
                pos.push_back(ll.y + y * distance + jy);
                pos.push_back(m_params.dims == 2 ? m_params.box_min[2] : ll.z + z * distance + jz);
                vel.insert(vel.end(), {0.0f, 0.0f, 0.0f});
                w.push_back(1.0f / mass);
                ro.push_back(1.0f);
                ph.push_back(PS_PHASE_SOLID);
            }
    addParticleMultiple(pos, vel, w, ro, ph);
}

// particlesystem.cpp:461-568: a horizontal cloth at height spacing.y with
// structural links (dist.x along x, dist.y along z); every x == 0 particle
// is pinned, and with holdEdges the other three edges too.
void ParticleSystem::addHorizCloth(Int2 ll, Int2 ur, Float3 spacing, Float2 dist, float mass,
                                   bool holdEdges)
{
    const uint32_t start = getNumParticles();
    const int cx = (int)((float)(int)std::ceil((float)(ur.x - ll.x)) / spacing.x);
    const int cz = (int)((float)(int)std::ceil((float)(ur.y - ll.y)) / spacing.z);
    std::vector<float> pos, vel, w, ro, pins, rest;
    std::vector<int32_t> ph;
    std::vector<uint32_t> pinIdx, pairs;
    for (int z = 0; z < cz; z++)
        for (int x = 0; x < cx; x++) {
            const float p[3] = {ll.x + x * spacing.x, spacing.y, ll.y + z * spacing.z};
            pos.insert(pos.end(), p, p + 3);
            const uint32_t idx = start + (uint32_t)(z * cx + x);
            auto pin = [&]() {
                pinIdx.push_back(idx);
                pins.insert(pins.end(), p, p + 3);
            };
            if (x > 0) {
                pairs.insert(pairs.end(), {idx - 1, idx});
                rest.push_back(dist.x);
            } else {
                pin();
            }
            if (z > 0) {
                pairs.insert(pairs.end(), {idx - (uint32_t)cx, idx});
                rest.push_back(dist.y);
            } else if (holdEdges) {
                pin();
            }
            if (x == cx - 1 && holdEdges) pin();
            if (z == cz - 1 && holdEdges) pin();
            vel.insert(vel.end(), {0.0f, 0.0f, 0.0f});
            w.push_back(1.0f / mass);
            ro.push_back(1.0f);
            ph.push_back(PS_PHASE_RIGID + m_rigidIndex);
        }
    if (addParticleMultiple(pos, vel, w, ro, ph) && !w.empty()) {
        if (!pinIdx.empty())
            check(ps_add_point_constraints(m_sys, (uint32_t)pinIdx.size(), pinIdx.data(), pins.data()));
        if (!rest.empty())
            check(ps_add_distance_constraints(m_sys, (uint32_t)rest.size(), pairs.data(), rest.data(),
                                              nullptr));
    }
    m_rigidIndex++;
}

// particlesystem.cpp:570-616
void ParticleSystem::addRope(Float3 start, Float3 spacing, float dist, int numLinks, float mass,
                             bool constrainStart)
{
    const uint32_t s0 = getNumParticles();
    std::vector<float> pos, vel, w, ro, rest;
    std::vector<int32_t> ph;
    std::vector<uint32_t> pairs;
    for (int i = 0; i <= numLinks; i++) {
        pos.insert(pos.end(), {start.x + i * spacing.x, start.y + i * spacing.y, start.z + i * spacing.z});
        vel.insert(vel.end(), {0.0f, 0.0f, 0.0f});
        w.push_back(1.0f / mass);
        ro.push_back(1.0f);
        ph.push_back(PS_PHASE_RIGID + m_rigidIndex);
        if (i > 0) {
            pairs.insert(pairs.end(), {s0 + i - 1, s0 + i});
            rest.push_back(dist);
        }
    }
    if (addParticleMultiple(pos, vel, w, ro, ph)) {
        if (!rest.empty())
            check(ps_add_distance_constraints(m_sys, (uint32_t)rest.size(), pairs.data(), rest.data(),
                                              nullptr));
        if (constrainStart) {
            const float p[3] = {start.x, start.y, start.z};
            check(ps_add_point_constraints(m_sys, 1, &s0, p));
        }
    }
    m_rigidIndex++;
}

// particlesystem.cpp:618-687: lattice points inside the inscribed sphere,
// inverse mass 0.01, each pinned where it starts
void ParticleSystem::addStaticSphere(Int3 ll, Int3 ur, float spacing)
{
    const uint32_t s0 = getNumParticles();
    const int cx = (int)((float)(int)std::ceil((float)(ur.x - ll.x)) / spacing);
    const int cy = (int)((float)(int)std::ceil((float)(ur.y - ll.y)) / spacing);
    const int cz = m_params.dims == 2 ? 1 : (int)((float)(int)std::ceil((float)(ur.z - ll.z)) / spacing);
    const float radius = (ur.x - ll.x) * 0.5f;
    const float c[3] = {ll.x + radius, ll.y + radius, ll.z + radius};
    std::vector<float> pos, vel, w, ro;
    std::vector<int32_t> ph;
    std::vector<uint32_t> idx;
    for (int z = 0; z < cz; z++)
        for (int y = 0; y < cy; y++)
            for (int x = 0; x < cx; x++) {
                const float p[3] = {ll.x + x * spacing, ll.y + y * spacing,
                                    m_params.dims == 2 ? m_params.box_min[2] : ll.z + z * spacing};
                const float dx = p[0] - c[0], dy = p[1] - c[1], dz = p[2] - c[2];
                if (std::sqrt(dx * dx + dy * dy + dz * dz) < radius) {
                    idx.push_back(s0 + (uint32_t)idx.size());
                    pos.insert(pos.end(), p, p + 3);
                    vel.insert(vel.end(), {0.0f, 0.0f, 0.0f});
                    w.push_back(0.01f);
                    ro.push_back(1.0f);
                    ph.push_back(PS_PHASE_RIGID + m_rigidIndex);
                }
            }
    if (addParticleMultiple(pos, vel, w, ro, ph) && !idx.empty())
        check(ps_add_point_constraints(m_sys, (uint32_t)idx.size(), idx.data(), pos.data()));
    m_rigidIndex++;
}

void ParticleSystem::setParticleToAdd(Float3 pos, Float3 vel, float mass)
{
    const float jitter = m_params.radius * 0.01f;  // particlesystem.cpp:258-261
    pos.x += (frand() * 2.0f - 1.0f) * jitter;
    pos.y += (frand() * 2.0f - 1.0f) * jitter;
    for (float v : {pos.x, pos.y, pos.z, vel.x, vel.y, vel.z, mass}) m_particlesToAdd.push_back(v);
}

void ParticleSystem::setFluidToAdd(Float3 pos, Float3 /*color*/, float mass, float density)
{
    for (float v : {pos.x, pos.y, pos.z, mass, density}) m_fluidsToAdd.push_back(v);
}

void ParticleSystem::makePointConstraint(unsigned index, Float3 point)
{
    const float p[3] = {point.x, point.y, point.z};
    check(ps_add_point_constraints(m_sys, 1, &index, p));
}

void ParticleSystem::makeDistanceConstraint(UInt2 index, float distance)
{
    const uint32_t pr[2] = {index.x, index.y};
    check(ps_add_distance_constraints(m_sys, 1, pr, &distance, nullptr));
}

void ParticleSystem::readPositions(std::vector<float> &xyz)
{
    xyz.resize(3 * (size_t)getNumParticles());
    check(ps_read_positions(m_sys, 0, getNumParticles(), xyz.data()));
}

void ParticleSystem::readVelocities(std::vector<float> &xyz)
{
    xyz.resize(3 * (size_t)getNumParticles());
    check(ps_read_velocities(m_sys, 0, getNumParticles(), xyz.data()));
}

ps_status ParticleSystem::status()
{
    ps_status s;
    check(ps_get_status(m_sys, &s));
    return s;
}

}  // namespace ps
