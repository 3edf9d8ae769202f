// ps_demo.cpp — drives the C++ ParticleSystem facade the way the reference's
// app does (gpu/src/particleapp.cpp): builds the reference's "combo" scene
// (key 8, particleapp.cpp:203-217) and the "ropes on a sphere" scene (key 9,
// :218-230), steps them at 60 Hz and prints a summary line per scene.
// Exit status 0 = every position finite and inside the box.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "particle_system.hpp"

using namespace ps;

static int run(const char *name, ParticleSystem &s, int frames)
{
    for (int f = 0; f < frames; f++) s.update(1.0f / 60.0f);
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

int main(int argc, char **argv)
{
    const int frames = argc > 1 ? std::atoi(argv[1]) : 120;
    int rc = 0;
    try {
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
        // a shot particle, queued like a mouse click (particleapp.cpp:140-145)
        s.setParticleToAdd(Float3{0, 30, 0}, Float3{0, -10, 0}, 2.f);
        rc |= run("combo", s, frames);

        // key 9: ropes hanging on an immovable sphere (particleapp.cpp:218-230)
        ParticleSystem t(0.25f, UInt3{64, 64, 64}, 15000, Int3{-50, 0, -50}, Int3{50, 200, 50}, 5);
        for (int i = 0; i < 50; i++) {
            const float ang = 3.14159265f * i * 0.02f;
            const Float3 v{std::cos(ang), std::sin(ang), 0.f};
            t.addRope(Float3{v.x * 5.f, v.y * 5.f + 10.f, 0.f}, Float3{v.x * .5f, v.y * .5f, 0.f}, .35f, 30, 1.f,
                      true);
        }
        t.addStaticSphere(Int3{-4, 7, -4}, Int3{4, 16, 4}, .5f);
        rc |= run("ropes_on_sphere", t, frames);
    } catch (const Error &e) {
        std::fprintf(stderr, "ps error %d: %s\n", e.code, e.what());
        return 2;
    }
    return rc;
}
