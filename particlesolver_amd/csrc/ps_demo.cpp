// ps_demo.cpp — drives the C++ ParticleSystem facade the way the reference's
// app does (gpu/src/particleapp.cpp): builds the reference's "combo" scene
// (key 8, particleapp.cpp:203-217) and the "ropes on a sphere" scene (key 9,
// :218-230) and ticks them at 60 Hz (ParticleApp::tick, :71-83).  In the
// combo run a particle is shot into the scene a third of the way through
// (left click: setParticleToAdd(eye, dir * 30, 2), :92-99) and the fluid
// emitter is switched on half way (space, :212-215: addFluid of a 2x1x2 box
// every 0.1 s).
//
// usage: ps_demo [FRAMES [DIR]]
// Prints a summary line per scene; exit status 0 = every position finite and
// inside the box.  With DIR, each scene's positions, velocities and phases are
// written to DIR/<scene>.pstraj before the first tick and after every tick.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "particle_system.hpp"

using namespace ps;

namespace {

const float kTick = 1.0f / 60.0f;

struct Recorder {
    ps_trajectory *traj = nullptr;
    Recorder(const char *dir, const char *name)
    {
        if (!dir) return;
        const std::string path = std::string(dir) + "/" + name + ".pstraj";
        if (ps_trajectory_open(path.c_str(), PS_FIELD_POSITIONS | PS_FIELD_VELOCITIES | PS_FIELD_PHASES, 3,
                               &traj) != PS_OK)
            throw Error(PS_ERR_INVALID_ARGUMENT, "cannot open " + path);
    }
    ~Recorder() { ps_trajectory_close(traj); }
    void write(ParticleSystem &s)
    {
        if (traj && ps_trajectory_write(traj, s.handle()) != PS_OK)
            throw Error(PS_ERR_INVALID_ARGUMENT, ps_last_error(s.handle()));
    }
};

int summary(const char *name, ParticleSystem &s, int frames)
{
    std::vector<float> x;
    s.readPositions(x);
    const Int3 lo = s.getMinBounds(), hi = s.getMaxBounds();
    double sy = 0.0;
    int bad = 0;
    for (size_t i = 0; i < x.size(); i += 3) {
        for (int a = 0; a < 3; a++)
            if (!std::isfinite(x[i + a])) bad++;
        if (x[i] < lo.x || x[i] > hi.x || x[i + 1] < lo.y || x[i + 1] > hi.y || x[i + 2] < lo.z ||
            x[i + 2] > hi.z)
            bad++;
        sy += x[i + 1];
    }
    const ps_status st = s.status();
    std::printf("%s: particles %u frames %d mean_y %.6f max_neighbors %u flags 0x%x bad %d\n", name,
                s.getNumParticles(), frames, x.empty() ? 0.0 : sy / (x.size() / 3),
                st.max_neighbors, st.flags, bad);
    return bad == 0 ? 0 : 1;
}

int combo(int frames, const char *dir)
{
    // key 8: combo scene (particleapp.cpp:203-217)
    ParticleSystem s(0.25f, UInt3{64, 64, 64}, 15000, Int3{-50, 0, -50}, Int3{50, 200, 50}, 5);
    s.addHorizCloth(Int2{14, -4}, Int2{24, 6}, Float3{.3f, 2.5f, .3f}, Float2{.25f, .25f}, 10.f, true);
    s.addHorizCloth(Int2{10, -10}, Int2{25, -5}, Float3{.3f, 15.5f, .3f}, Float2{.25f, .25f}, 3.f, false);
    s.addRope(Float3{-17, 20, -17}, Float3{0, -.5f, 0.001f}, .4f, 30, 1.f, true);
    s.addRope(Float3{-16, 20, -17}, Float3{0, 0, .5f}, .4f, 50, 1.f, true);
    s.addRope(Float3{-17, 20, -16}, Float3{0, -.5f, 0.001f}, .4f, 40, 1.f, true);
    s.addParticleGrid(Int3{17, 6, 0}, Int3{21, 11, 4}, 1.f, false);
    s.addParticleGrid(Int3{-12, 0, -20}, Int3{0, 12, -17}, 1.f, false);
    s.addParticleGrid(Int3{-18, 0, -15}, Int3{-16, 9, -12}, 1.f, false);
    s.addStaticSphere(Int3{5, 5, -10}, Int3{10, 10, -5}, .5f);
    Recorder rec(dir, "combo");
    rec.write(s);
    bool emitter = false;
    float timer = -1.0f;
    for (int f = 0; f < frames; f++) {
        if (f == frames / 3)  // a click: eye (0, 25, 20) looking along (0, -0.6, -0.8)
            s.setParticleToAdd(Float3{0.f, 25.f, 20.f}, Float3{0.f, -0.6f * 30.f, -0.8f * 30.f}, 2.f);
        if (f == frames / 2) emitter = true;
        if (emitter && timer <= 0.0f) {
            s.addFluid(Int3{-1, 0, -1}, Int3{1, 1, 1}, 1.f, 1.f, Float3{0, 0, 1});
            timer = 0.1f;
        }
        timer -= kTick;
        s.update(kTick);
        rec.write(s);
    }
    return summary("combo", s, frames);
}

int ropes_on_sphere(int frames, const char *dir)
{
    // key 9: fifty ropes hung from a ring around an immovable sphere
    // (particleapp.cpp:218-230); the angle is the reference's float
    // M_PI * i * 0.02f, its cosine and sine are taken in double and rounded
    ParticleSystem t(0.25f, UInt3{64, 64, 64}, 15000, Int3{-50, 0, -50}, Int3{50, 200, 50}, 5);
    for (int i = 0; i < 50; i++) {
        const float ang = (float)(M_PI * i * 0.02f);
        const Float3 v{(float)std::cos((double)ang), (float)std::sin((double)ang), 0.f};
        t.addRope(Float3{v.x * 5.f + 0.f, v.y * 5.f + 10.f, v.z * 5.f + 0.f},
                  Float3{v.x * .5f, v.y * .5f, v.z * .5f}, .35f, 30, 1.f, true);
    }
    t.addStaticSphere(Int3{-4, 7, -4}, Int3{4, 16, 4}, .5f);
    Recorder rec(dir, "ropes_on_sphere");
    rec.write(t);
    for (int f = 0; f < frames; f++) {
        t.update(kTick);
        rec.write(t);
    }
    return summary("ropes_on_sphere", t, frames);
}

}  // namespace

int main(int argc, char **argv)
{
    const int frames = argc > 1 ? std::atoi(argv[1]) : 120;
    const char *dir = argc > 2 ? argv[2] : nullptr;
    int rc = 0;
    try {
        rc |= combo(frames, dir);
        rc |= ropes_on_sphere(frames, dir);
    } catch (const Error &e) {
        std::fprintf(stderr, "ps error %d: %s\n", e.code, e.what());
        return 2;
    }
    return rc;
}
