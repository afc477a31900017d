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
#include <map>
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

// The reference's test systems (api/nblib/tests/testsystems.cpp:50-335),
// built the way it builds them: a particle library, Molecule with
// exclusions by name, TopologyBuilder.
struct ParticleLibrary
{
    // testsystems.cpp:53-82
    std::map<std::string, ParticleType> types{
        { "Ow", ParticleType(ParticleTypeName("Ow"), Mass(15.99940f)) },
        { "H", ParticleType(ParticleTypeName("H"), Mass(1.008f)) },
        { "OMet", ParticleType(ParticleTypeName("OMet"), Mass(15.999f)) },
        { "CMet", ParticleType(ParticleTypeName("CMet"), Mass(15.035f)) },
        { "Ar", ParticleType(ParticleTypeName("Ar"), Mass(39.94800f)) },
    };
    std::map<std::string, real> c6{ { "Ow", 0.0026173456f }, { "H", 0.0f }, { "OMet", 0.0022619536f },
                                    { "CMet", 0.0088755241f }, { "Ar", 0.0062647225f } };
    std::map<std::string, real> c12{ { "Ow", 2.634129e-06f }, { "H", 0.0f }, { "OMet", 1.505529e-06f },
                                     { "CMet", 2.0852922e-05f }, { "Ar", 9.847044e-06f } };
    const ParticleType& type(const std::string& n) const { return types.at(n); }
    ParticleTypesInteractions interactions(const std::vector<std::string>& names) const
    {
        ParticleTypesInteractions p;
        for (const std::string& n : names)
        {
            p.add(ParticleTypeName(n), C6(c6.at(n)), C12(c12.at(n)));
        }
        return p;
    }
};

// testsystems.cpp:104-121 and 130-160
Molecule waterMolecule(const ParticleLibrary& plib)
{
    Molecule water(MoleculeName("SOL"));
    water.addParticle(ParticleName("Oxygen"), Charge(-0.82f), plib.type("Ow"));
    water.addParticle(ParticleName("H1"), Charge(+0.41f), plib.type("H"));
    water.addParticle(ParticleName("H2"), Charge(+0.41f), plib.type("H"));
    water.addExclusion("H1", "Oxygen");
    water.addExclusion("H2", "Oxygen");
    water.addExclusion("H1", "H2");
    return water;
}

Molecule methanolMolecule(const ParticleLibrary& plib)
{
    Molecule methanol(MoleculeName("MeOH"));
    methanol.addParticle(ParticleName("Me1"), Charge(+0.176f), plib.type("CMet"));
    methanol.addParticle(ParticleName("O2"), Charge(-0.574f), plib.type("OMet"));
    methanol.addParticle(ParticleName("H3"), Charge(+0.398f), plib.type("H"));
    methanol.addExclusion("Me1", "O2");
    methanol.addExclusion("Me1", "H3");
    methanol.addExclusion("H3", "O2");
    return methanol;
}

// SpcMethanolTopologyBuilder::buildTopology (testsystems.cpp:188-208)
Topology spcMethanolTopology(int numWater, int numMethanol)
{
    ParticleLibrary plib;
    TopologyBuilder builder;
    builder.addMolecule(methanolMolecule(plib), numMethanol);
    builder.addMolecule(waterMolecule(plib), numWater);
    builder.addParticleTypesInteractions(plib.interactions({ "Ow", "H", "OMet", "CMet" }));
    return builder.buildTopology();
}

// ArgonTopologyBuilder (testsystems.cpp:225-245)
Topology argonTopology(int numParticles)
{
    ParticleLibrary plib;
    Molecule        argon(MoleculeName("AR"));
    argon.addParticle(ParticleName("AR"), plib.type("Ar"));
    TopologyBuilder builder;
    builder.addMolecule(argon, numParticles);
    builder.addParticleTypesInteractions(plib.interactions({ "Ar" }));
    return builder.buildTopology();
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
    run("TopologyBuilderSpcMethanol", [] {
        // the layout the reference's builder produces: methanol then water,
        // charges per particle, each particle excluding its molecule
        const Topology t = spcMethanolTopology(1, 1);
        expect(t.numParticles() == 6, "6 particles");
        const std::vector<real> q{ 0.176f, -0.574f, 0.398f, -0.82f, 0.41f, 0.41f };
        expect(t.getCharges() == q, "charges");
        const ExclusionLists& ex = t.getGmxExclusions();
        const std::vector<int> off{ 0, 3, 6, 9, 12, 15, 18 };
        const std::vector<int> idx{ 0, 1, 2, 0, 1, 2, 0, 1, 2, 3, 4, 5, 3, 4, 5, 3, 4, 5 };
        expect(ex.offsets == off && ex.indices == idx, "exclusions");
        const std::vector<std::string> names{ "CMet", "OMet", "H", "Ow", "H", "H" };
        for (int a = 0; a < 6; a++)
        {
            const std::string& n = t.getParticleTypes()[t.getParticleTypeIdOfAllParticles()[a]].name();
            expect(n == names[a], "type of particle " + std::to_string(a) + " is " + n);
        }
        expect(t.getParticleTypes().size() == 4, "4 distinct types");
        // geometric combination of the library's C6 (interactions.cpp:83-93)
        const real c6 = t.getNonBondedInteractionMap().getC6("Ow", "CMet");
        expect(std::fabs(c6 - std::sqrt(0.0026173456f * 0.0088755241f)) <= 1e-9f, "combined C6");
        expect(t.sequenceID(MoleculeName("SOL"), 0, ResidueName("SOL"), ParticleName("H2")) == 5, "sequenceID");
        expect(t.sequenceID(MoleculeName("MeOH"), 0, ResidueName("MeOH"), ParticleName("O2")) == 1, "sequenceID");
        const Topology two = spcMethanolTopology(2, 1);
        expect(two.numParticles() == 9 && two.getGmxExclusions().indices[21] == 6, "second water offset");
        expect(two.sequenceID(MoleculeName("SOL"), 1, ResidueName("SOL"), ParticleName("Oxygen")) == 6,
               "second water");
    });
    run("TopologyBuilderArgon", [] {
        const Topology t = argonTopology(12);
        expect(t.numParticles() == 12 && t.getParticleTypes().size() == 1, "12 argon atoms of one type");
        for (int a = 0; a < 12; a++)
        {
            const ExclusionLists& ex = t.getGmxExclusions();
            expect(ex.offsets[a + 1] - ex.offsets[a] == 1 && ex.indices[ex.offsets[a]] == a, "self exclusion");
            expect(t.getCharges()[a] == 0.0f, "charge 0");
        }
    });
    run("MoleculeInputErrors", [] {
        ParticleLibrary plib;
        Molecule        m(MoleculeName("X"));
        m.addParticle(ParticleName("A"), plib.type("Ow"));
        // same type name, different mass
        expectThrowsInput([&] { m.addParticle(ParticleName("B"), ParticleType(ParticleTypeName("Ow"), Mass(1.0f))); },
                          "differing particle types");
        m.addExclusion("A", "Nope");
        expectThrowsInput([&] { m.getExclusions(); }, "exclusion by unknown name");
        expectThrowsInput([&] { m.addExclusion(0, 3); }, "exclusion index out of range");
        // a type without parameters
        TopologyBuilder b;
        Molecule        w = waterMolecule(plib);
        b.addMolecule(w, 2);
        b.addParticleTypesInteractions(plib.interactions({ "Ow" }));
        expectThrowsInput([&] { b.buildTopology(); }, "missing interaction parameters");
        TopologyBuilder b2;
        Molecule        other(MoleculeName("Y"));
        other.addParticle(ParticleName("O"), ParticleType(ParticleTypeName("Ow"), Mass(2.0f)));
        b2.addMolecule(w, 1);
        expectThrowsInput([&] { b2.addMolecule(other, 1); }, "differing types across molecules");
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
    // the same systems with the topology from Molecule/TopologyBuilder
    // (testsystems.cpp:244-335) instead of the arrays
    for (const TestSystem& sys : systems)
    {
        run(sys.name + "BuiltWithTopologyBuilderForcesAreCorrect", [&] {
            const Topology top = sys.name == "Argon" ? argonTopology(static_cast<int>(sys.x.size()))
                                                     : spcMethanolTopology(1, 1);
            std::vector<Vec3> zeros(sys.x.size());
            SimulationState   state(sys.x, zeros, zeros, Box(sys.box), top);
            ForceCalculator   calc(state, plainCutoff(sys.rc));
            ArrayRef<Vec3>    forces(state.forces());
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
                           "force " + std::to_string(i) + "." + std::to_string(d));
                }
            }
        });
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
