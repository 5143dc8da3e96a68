// oracle_rng.cpp -- the reference's Langevin noise stream (test infrastructure
// only; part of the checker, never of the product).
//
// celeste::random::Random (random/Random.cpp:10-43) wraps a std::mt19937
// seeded with the run's --random-seed (MmSystem::set_random, MmSystem.cpp:818-821)
// and draws each normal from a freshly constructed std::normal_distribution
// (Random.cpp:40-43), so the value cached by the distribution is discarded
// on every draw.  SubBox::update_coordinates_langevin (SubBox.cpp:2077-2114)
// draws three per atom per step, in atom order, on every step after the
// first.  This file calls the same libstdc++ generator and distribution in
// the same way; it restates no algorithm of its own.
#include <random>

extern "C" void orc_ref_normals(int seed, long long skip, long long n, double *out) {
    std::mt19937 gen;
    gen.seed(seed);
    for (long long k = 0; k < skip; k++) {
        std::normal_distribution<> d(0.0, 1.0);
        (void)d(gen);
    }
    for (long long k = 0; k < n; k++) {
        std::normal_distribution<> d(0.0, 1.0);
        out[k] = d(gen);
    }
}
