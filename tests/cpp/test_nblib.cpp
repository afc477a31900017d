// Tests of the nblib-compatible C++ layer, mirroring the reference's
// api/nblib/tests/nbkernelsystem.cpp (ForceCalculator on the SPC-methanol
// and argon systems of testsystems.cpp) plus host-side input validation
// (box.cpp, simulationstate.cpp, interactions.cpp, topology checks).
//
// Usage: test_nblib host                    host-only checks (no device)
//        test_nblib device <system file>... force tests on the GPU
// A system file (written by tests/test_nblib_cpp.py from the golden
// fixtures) holds: name; natoms ntypes; box rc; ntypes lines "name mass c6
// c12"; natoms lines "type charge x y z"; natoms exclusion lines "n idx...";
// natoms lines of reference forces.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <functional>
#include <iostream>
#include <string>
#include <vector>

#include "nblib.h"

using namespace nblib;

namespace
{

int g_failures = 0;

void run(const std::string& name, const std::function<void()>& body)
{
    try
    {
        body();
        std::printf("PASS %s\n", name.c_str());
    }
    catch (const std::exception& e)
    {
        g_failures++;
        std::printf("FAIL %s: %s\n", name.c_str(), e.what());
    }
    std::fflush(stdout);
}

void expect(bool ok, const std::string& what)
{
    if (!ok)
    {
        throw std::runtime_error(what);
    }
}

template<typename F>
void expectThrowsInput(F&& f, const std::string& what)
{
    try
    {
        f();
    }
    catch (const InputException&)
    {
        return;
    }
    throw std::runtime_error("expected InputException: " + what);
}

struct TestSystem
{
    std::string       name;
    real              box = 0, rc = 1;
    std::vector<ParticleType> types;
    std::vector<real> c6, c12;
    std::vector<int>  typeIds;
    std::vector<real> charges;
    std::vector<Vec3> x, fRef;
    ExclusionLists    excl;

    SimulationState state() const
    {
        ParticleTypesInteractions interactions;
        for (size_t t = 0; t < types.size(); t++)
        {
            interactions.add(types[t].name(), c6[t], c12[t]);
        }
        Topology top(types, typeIds, charges, excl, interactions.generateTable());
        std::vector<Vec3> zeros(x.size());
        return SimulationState(x, zeros, zeros, Box(box), top);
    }
};

TestSystem load(const std::string& path)
{
    std::ifstream in(path);
    if (!in)
    {
        throw std::runtime_error("cannot read " + path);
    }
    TestSystem s;
    int        n = 0, nt = 0;
    in >> s.name >> n >> nt >> s.box >> s.rc;
    for (int t = 0; t < nt; t++)
    {
        std::string name;
        real        mass, c6, c12;
        in >> name >> mass >> c6 >> c12;
        s.types.emplace_back(name, mass);
        s.c6.push_back(c6);
        s.c12.push_back(c12);
    }
    for (int a = 0; a < n; a++)
    {
        int  t;
        real q, x, y, z;
        in >> t >> q >> x >> y >> z;
        s.typeIds.push_back(t);
        s.charges.push_back(q);
        s.x.emplace_back(x, y, z);
    }
    for (int a = 0; a < n; a++)
    {
        int k;
        in >> k;
        for (int i = 0; i < k; i++)
        {
            int j;
            in >> j;
            s.excl.indices.push_back(j);
        }
        s.excl.offsets.push_back(static_cast<int>(s.excl.indices.size()));
    }
    for (int a = 0; a < n; a++)
    {
        real fx, fy, fz;
        in >> fx >> fy >> fz;
        s.fRef.emplace_back(fx, fy, fz);
    }
    if (!in)
    {
        throw std::runtime_error("malformed system file " + path);
    }
    return s;
}

NBKernelOptions plainCutoff(real rc)
{
    NBKernelOptions o;
    o.nbnxmSimd      = SimdKernels::SimdNo;
    o.coulombType    = CoulombType::Cutoff;
    o.pairlistCutoff = rc;
    return o;
}

void hostTests()
{
    run("BoxRejectsNaN", [] { expectThrowsInput([] { Box b(std::nanf("")); }, "NaN box"); });
    run("BoxRejectsInf", [] { expectThrowsInput([] { Box b(1.0f, INFINITY, 1.0f); }, "Inf box"); });
    run("ParticleTypesInteractionsRejectsConflictingDuplicate", [] {
        ParticleTypesInteractions p;
        p.add("Ow", 1e-3f, 1e-6f);
        p.add("Ow", 1e-3f, 1e-6f); // identical duplicate is accepted
        expectThrowsInput([&] { p.add("Ow", 2e-3f, 1e-6f); }, "conflicting duplicate");
    });
    run("LorentzBerthelotIsRejected", [] {
        ParticleTypesInteractions p(CombinationRule::LorentzBerthelot);
        p.add("Ar", 1e-3f, 1e-6f);
        expectThrowsInput([&] { p.generateTable(); }, "LB rule");
    });
    run("GeometricCombination", [] {
        ParticleTypesInteractions p;
        p.add("A", 4.0f, 9.0f).add("B", 1.0f, 1.0f).add("A", "B", 0.5f, 0.25f);
        const NonBondedInteractionMap m = p.generateTable();
        expect(m.getC6("A", "A") == 4.0f && m.getC12("A", "A") == 9.0f, "self pair");
        expect(m.getC6("A", "B") == 0.5f && m.getC12("B", "A") == 0.25f, "override wins");
    });
    run("TopologyRejectsBadTypeId", [] {
        ParticleTypesInteractions p;
        p.add("A", 1e-3f, 1e-6f);
        ExclusionLists ex;
        ex.indices = { 0 };
        ex.offsets = { 0, 1 };
        expectThrowsInput([&] { Topology t({ ParticleType("A", 1) }, { 1 }, { 0.0f }, ex, p.generateTable()); },
                          "type id");
    });
    run("SimulationStateRejectsNaN", [] {
        ParticleTypesInteractions p;
        p.add("A", 1e-3f, 1e-6f);
        ExclusionLists ex;
        ex.indices = { 0 };
        ex.offsets = { 0, 1 };
        Topology          t({ ParticleType("A", 1) }, { 0 }, { 0.0f }, ex, p.generateTable());
        std::vector<Vec3> x{ Vec3(std::nanf(""), 0, 0) }, z(1);
        expectThrowsInput([&] { SimulationState s(x, z, z, Box(2.0f), t); }, "NaN coordinate");
    });
}

void forcesAreCorrect(const TestSystem& sys)
{
    // nbkernelsystem.cpp:69-85,187-203
    SimulationState state = sys.state();
    ForceCalculator calc(state, plainCutoff(sys.rc));
    ArrayRef<Vec3>  forces(state.forces());
    calc.compute(state.coordinates(), forces);
    real fmax = 0;
    for (const Vec3& f : sys.fRef)
    {
        for (int d = 0; d < 3; d++)
        {
            fmax = std::max(fmax, std::fabs(f[d]));
        }
    }
    for (size_t i = 0; i < forces.size(); i++)
    {
        for (int d = 0; d < 3; d++)
        {
            expect(std::fabs(forces[i][d] - sys.fRef[i][d]) <= 5e-5f * fmax,
                   "force " + std::to_string(i) + "." + std::to_string(d) + " " + std::to_string(forces[i][d])
                           + " vs " + std::to_string(sys.fRef[i][d]));
        }
    }
}

void deviceTests(const std::vector<TestSystem>& systems)
{
    for (const TestSystem& sys : systems)
    {
        run(sys.name + "ForcesAreCorrect", [&] { forcesAreCorrect(sys); });
    }
    const TestSystem& spc = systems.back();
    run("ExpectedNumberOfForces", [&] {
        // nbkernelsystem.cpp:87-101
        SimulationState state = spc.state();
        ForceCalculator calc(state, plainCutoff(spc.rc));
        ArrayRef<Vec3>  forces(state.forces());
        calc.compute(state.coordinates(), forces);
        expect(static_cast<int>(forces.size()) == state.topology().numParticles(), "one force per particle");
    });
    run("CanIntegrateSystem", [&] {
        // nbkernelsystem.cpp:103-122
        SimulationState state = spc.state();
        ForceCalculator calc(state, plainCutoff(spc.rc));
        LeapFrog        integrator(state.topology(), state.box());
        calc.compute(state.coordinates(), state.forces());
        integrator.integrate(1.0f, state.coordinates(), state.velocities(), state.forces());
    });
    run("UpdateChangesForces", [&] {
        // nbkernelsystem.cpp:124-184.  Repeated computes agree to rounding
        // (float atomics order the sums differently), not bit for bit.
        SimulationState state = spc.state();
        ForceCalculator calc(state, plainCutoff(spc.rc));
        LeapFrog        integrator(state.topology(), state.box());
        calc.compute(state.coordinates(), state.forces());
        const std::vector<Vec3> f1 = state.forces();
        calc.compute(state.coordinates(), state.forces());
        for (size_t i = 0; i < f1.size(); i++)
        {
            for (int d = 0; d < 3; d++)
            {
                expect(std::fabs(state.forces()[i][d] - f1[i][d]) <= 1e-5f * (1.0f + std::fabs(f1[i][d])),
                       "forces changed without an update");
            }
        }
        integrator.integrate(1.0f, state.coordinates(), state.velocities(), state.forces());
        calc.compute(state.coordinates(), state.forces());
        for (size_t i = 0; i < f1.size(); i++)
        {
            for (int d = 0; d < 3; d++)
            {
                expect(state.forces()[i][d] != f1[i][d], "forces unchanged after an update");
            }
        }
    });
    run("EnergiesAndVirialWhenRequested", [&] {
        SimulationState state = spc.state();
        NBKernelOptions o     = plainCutoff(spc.rc);
        o.computeVirialAndEnergy = true;
        ForceCalculator calc(state, o);
        calc.compute(state.coordinates(), state.forces());
        const auto e = calc.energies();
        const auto v = calc.virial();
        expect(std::isfinite(e[0]) && std::isfinite(e[1]) && e[1] != 0.0, "energies");
        expect(std::isfinite(v[0]) && std::isfinite(v[4]) && std::isfinite(v[8]), "virial");
    });
    run("ComputeRejectsWrongSizes", [&] {
        SimulationState   state = spc.state();
        ForceCalculator   calc(state, plainCutoff(spc.rc));
        std::vector<Vec3> small(2);
        expectThrowsInput([&] { calc.compute(state.coordinates(), small); }, "force array size");
    });
    run("AnalyticalEwaldDefaultOptions", [&] {
        // nblib's default options: Pme with the analytical correction
        SimulationState state = spc.state();
        NBKernelOptions o;
        o.pairlistCutoff = spc.rc;
        ForceCalculator calc(state, o);
        calc.compute(state.coordinates(), state.forces());
        for (const Vec3& f : state.forces())
        {
            expect(std::isfinite(f[0]) && std::isfinite(f[1]) && std::isfinite(f[2]), "finite forces");
        }
    });
}

} // namespace

int main(int argc, char** argv)
{
    if (argc < 2)
    {
        std::fprintf(stderr, "usage: %s host | device <system file>...\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    if (mode == "host")
    {
        hostTests();
    }
    else if (mode == "device" && argc > 2)
    {
        std::vector<TestSystem> systems;
        for (int i = 2; i < argc; i++)
        {
            systems.push_back(load(argv[i]));
        }
        deviceTests(systems);
    }
    else
    {
        std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
        return 2;
    }
    std::printf("%d failure(s)\n", g_failures);
    return g_failures == 0 ? 0 : 1;
}
