// particle_system.hpp — C++ host API with the reference's ParticleSystem
// surface (gpu/src/particlesystem.h:22-118), implemented over the C-ABI in
// include/ps.h.  Scene builders follow gpu/src/particlesystem.cpp:357-698.
//
// Differences a caller of the reference sees:
//  - positions are read with readPositions() (the reference keeps them in a
//    GL vertex buffer, getCurrentReadBuffer, particlesystem.h:47); rendering
//    metadata (colours) is not kept;
//  - the grid is dense, with cell = max(h, 2r) and sized to the bounds, instead
//    of a wrapped 64^3 hash (integration_kernel.cuh:199-201);
//  - errors throw ps::Error instead of printing and exiting; additions past
//    maxParticles are dropped silently as there (particlesystem.cpp:303,335),
//    and a builder whose particles were dropped adds no constraints;
//  - random jitter comes from a seeded SplitMix64 stream instead of rand().
#pragma once

#include <cstdint>
#include <deque>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ps.h"

namespace ps {

struct Int2 { int x, y; };
struct Int3 { int x, y, z; };
struct UInt2 { unsigned x, y; };
struct UInt3 { unsigned x, y, z; };
struct Float2 { float x, y; };
struct Float3 { float x, y, z; };

class Error : public std::runtime_error {
public:
    Error(int code, const std::string &msg) : std::runtime_error(msg), code(code) {}
    int code;
};

class ParticleSystem {
public:
    // particlesystem.h:25.  gridSize is accepted for signature compatibility;
    // the dense grid is derived from the bounds and the kernel radius h.
    ParticleSystem(float particleRadius, UInt3 gridSize, unsigned maxParticles, Int3 minBounds,
                   Int3 maxBounds, int iterations, float h = 2.0f);
    // Construct from a full parameter set (no reference counterpart).
    explicit ParticleSystem(const ps_params &params);
    ~ParticleSystem();
    ParticleSystem(const ParticleSystem &) = delete;
    ParticleSystem &operator=(const ParticleSystem &) = delete;

    // particlesystem.cpp:144-246 (dt clamped to 0.05 as there, :149), then
    // the deferred additions (addNewStuff, :249-253)
    void update(float deltaTime);

    // scene builders, particlesystem.cpp:357-687
    void addFluid(Int3 ll, Int3 ur, float mass, float density, Float3 color = {0, 0, 1});
    void addParticleGrid(Int3 ll, Int3 ur, float mass, bool addJitter);
    void addHorizCloth(Int2 ll, Int2 ur, Float3 spacing, Float2 dist, float mass, bool holdEdges);
    void addRope(Float3 start, Float3 spacing, float dist, int numLinks, float mass,
                 bool constrainStart);
    void addStaticSphere(Int3 ll, Int3 ur, float spacing);

    // deferred single-particle / fluid additions (particlesystem.cpp:256-331)
    void setParticleToAdd(Float3 pos, Float3 vel, float mass);
    void setFluidToAdd(Float3 pos, Float3 color, float mass, float density);

    // particlesystem.cpp:690-698
    void makePointConstraint(unsigned index, Float3 point);
    void makeDistanceConstraint(UInt2 index, float distance);

    unsigned getNumParticles() const;
    float getParticleRadius() const { return m_params.radius; }
    Int3 getMinBounds() const { return m_minBounds; }
    Int3 getMaxBounds() const { return m_maxBounds; }

    // positions / velocities in insertion order, 3 floats per particle
    void readPositions(std::vector<float> &xyz);
    void readVelocities(std::vector<float> &xyz);
    ps_status status();
    ps_system *handle() { return m_sys; }

private:
    void check(int rc) const;
    bool addParticleMultiple(const std::vector<float> &pos, const std::vector<float> &vel,
                             const std::vector<float> &w, const std::vector<float> &ro,
                             const std::vector<int32_t> &phase, bool single = false);
    float frand();

    ps_system *m_sys = nullptr;
    ps_params m_params;
    Int3 m_minBounds{0, 0, 0}, m_maxBounds{0, 0, 0};
    int m_rigidIndex = 0;
    uint64_t m_rng = 0;
    std::deque<float> m_particlesToAdd;  // 7 floats each: pos, vel, mass
    std::deque<float> m_fluidsToAdd;     // 5 floats each: pos, mass, density
};

}  // namespace ps
