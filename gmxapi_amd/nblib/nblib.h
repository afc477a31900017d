// nblib-compatible C++ host API over the MI355X nonbonded engine (nbx.h).
//
// Mirrors the public surface of the reference's nblib that drives the NBNxM
// nonbonded path: Box (api/nblib/box.h:60-79), NBKernelOptions and its enums
// (kerneloptions.h:50-112), ParticleType and NonBondedInteractionMap
// (particletype.h, interactions.h:69-120), Topology (topology.h:72-121),
// SimulationState (simulationstate.h:66-97), ForceCalculator
// (forcecalculator.h:60-106) and LeapFrog (integrator.h).  Names, argument
// meaning and errors follow the reference: invalid input throws
// nblib::InputException.  The force backend is the HIP library; there is no
// CPU fallback, so constructing a ForceCalculator without a device throws.
//
// Topologies are built as in the reference, through Molecule
// (molecules.h:80-170) and TopologyBuilder (topology.h:133-175) with the
// strong types of the reference's call sites (ParticleName, Charge,
// MoleculeName, ParticleTypeName, Mass, C6, C12); a Topology can also be
// made directly from the arrays the builder produces.  Type ids are given
// in order of first appearance over the added molecules (the reference
// numbers them in its hash map's iteration order, which only permutes the
// ids).  Other differences, all additive: ForceCalculator exposes energies
// and the virial of the last compute, and NBKernelOptions carries the HIP
// device ordinal.
#pragma once

#include <array>
#include <cstddef>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

namespace nblib
{

using real = float;

//! xyz triplet (gmx::RVec)
struct Vec3
{
    real v[3] = { 0, 0, 0 };
    Vec3() = default;
    Vec3(real x, real y, real z) : v{ x, y, z } {}
    real&       operator[](int d) { return v[d]; }
    const real& operator[](int d) const { return v[d]; }
};

//! Non-owning view of a contiguous array (gmx::ArrayRef)
template<typename T>
class ArrayRef
{
public:
    ArrayRef() = default;
    ArrayRef(T* data, std::size_t size) : data_(data), size_(size) {}
    template<typename C>
    ArrayRef(C& c) : data_(c.data()), size_(c.size())
    {
    }
    T*          data() const { return data_; }
    std::size_t size() const { return size_; }
    T&          operator[](std::size_t i) const { return data_[i]; }
    T*          begin() const { return data_; }
    T*          end() const { return data_ + size_; }

private:
    T*          data_ = nullptr;
    std::size_t size_ = 0;
};

//! A value of type T tagged by Phantom, explicit to construct
//! (nblib/util/user.h's StrongType): names and quantities cannot be mixed up
template<typename T, typename Phantom>
class StrongType
{
public:
    StrongType() = default;
    template<typename U>
    explicit StrongType(U&& v) : value_(std::forward<U>(v))
    {
    }
    const T& value() const { return value_; }
    operator const T&() const { return value_; }
    bool operator==(const StrongType& o) const { return value_ == o.value_; }
    bool operator<(const StrongType& o) const { return value_ < o.value_; }

private:
    T value_{};
};

using ParticleTypeName = StrongType<std::string, struct ParticleTypeNameParameter>;
using Mass             = StrongType<real, struct MassParameter>;
using C6               = StrongType<real, struct C6Parameter>;
using C12              = StrongType<real, struct C12Parameter>;
using ParticleName     = StrongType<std::string, struct ParticleNameParameter>;
using Charge           = StrongType<real, struct ChargeParameter>;
using ResidueName      = StrongType<std::string, struct ResidueNameParameter>;
using MoleculeName     = StrongType<std::string, struct MoleculeNameParameter>;

//! Invalid user input (api/nblib/exception.h)
class InputException : public std::runtime_error
{
public:
    explicit InputException(const std::string& what) : std::runtime_error(what) {}
};

//! Cubic or rectangular periodic box (box.h:60-79)
class Box
{
public:
    using LegacyMatrix = std::array<std::array<real, 3>, 3>;
    explicit Box(real l);
    Box(real x, real y, real z);
    const LegacyMatrix& legacyMatrix() const { return m_; }

private:
    LegacyMatrix m_{};
};

enum class SimdKernels : int
{
    SimdAuto,
    SimdNo,
    Simd4XM,
    Simd2XMM,
    Count
};

enum class CombinationRule : int
{
    Geometric,
    LorentzBerthelot,
    None,
    Count
};

enum class CoulombType : int
{
    Pme,
    Cutoff,
    ReactionField,
    Count
};

//! kerneloptions.h:86-112
struct NBKernelOptions
{
    bool            useGpu                 = true;
    int             numOpenMPThreads       = 1;
    SimdKernels     nbnxmSimd              = SimdKernels::SimdAuto;
    CombinationRule ljCombinationRule      = CombinationRule::Geometric;
    bool            useHalfLJOptimization  = false;
    real            pairlistCutoff         = 1.0;
    bool            computeVirialAndEnergy = false;
    CoulombType     coulombType            = CoulombType::Pme;
    bool            useTabulatedEwaldCorr  = false;
    int             numIterations          = 100;
    bool            cyclesPerPair          = false;
    real            timestep               = 0.001;
    //! HIP device ordinal (extension)
    int device = 0;
};

//! particletype.h: a named type with a mass
class ParticleType
{
public:
    ParticleType(ParticleTypeName name, Mass mass) : name_(name.value()), mass_(mass.value()) {}
    ParticleType(std::string name, real mass) : name_(std::move(name)), mass_(mass) {}
    const std::string& name() const { return name_; }
    real               mass() const { return mass_; }
    //! particletype.cpp: same name and mass
    bool operator==(const ParticleType& o) const { return name_ == o.name_ && mass_ == o.mass_; }

private:
    std::string name_;
    real        mass_;
};

//! interactions.h:69-91: C6/C12 per pair of type names
class NonBondedInteractionMap
{
public:
    void setInteractions(const std::string& a, const std::string& b, real c6, real c12);
    real getC6(const std::string& a, const std::string& b) const;
    real getC12(const std::string& a, const std::string& b) const;

private:
    std::map<std::pair<std::string, std::string>, std::pair<real, real>> map_;
};

//! interactions.h:100-140: per-type parameters combined by a rule, with
//! optional per-pair overrides
class ParticleTypesInteractions
{
public:
    explicit ParticleTypesInteractions(CombinationRule rule = CombinationRule::Geometric) : rule_(rule) {}
    ParticleTypesInteractions& add(const std::string& name, real c6, real c12);
    ParticleTypesInteractions& add(const std::string& a, const std::string& b, real c6, real c12);
    //! the reference's signatures (interactions.h:110-118)
    ParticleTypesInteractions& add(const ParticleTypeName& name, C6 c6, C12 c12)
    {
        return add(name.value(), c6.value(), c12.value());
    }
    ParticleTypesInteractions& add(const ParticleTypeName& a, const ParticleTypeName& b, C6 c6, C12 c12)
    {
        return add(a.value(), b.value(), c6.value(), c12.value());
    }
    NonBondedInteractionMap generateTable() const;
    CombinationRule         getCombinationRule() const { return rule_; }
    //! interactions.cpp:192-220: the other's parameters join these (same rule;
    //! conflicting duplicates throw)
    void merge(const ParticleTypesInteractions& other);

private:
    CombinationRule                                                      rule_;
    std::map<std::string, std::pair<real, real>>                         single_;
    std::map<std::pair<std::string, std::string>, std::pair<real, real>> pairs_;
};

//! gmx::ListOfLists<int> in CSR form
struct ExclusionLists
{
    std::vector<int> offsets{ 0 };
    std::vector<int> indices;
    int              numLists() const { return static_cast<int>(offsets.size()) - 1; }
};

//! topology.h:72-121, built from the arrays TopologyBuilder would produce
class Topology
{
public:
    Topology(std::vector<ParticleType> particleTypes, std::vector<int> typeIdOfAllParticles,
             std::vector<real> charges, ExclusionLists exclusions, NonBondedInteractionMap interactions,
             CombinationRule rule = CombinationRule::Geometric);

    int                            numParticles() const { return static_cast<int>(typeIds_.size()); }
    const std::vector<ParticleType>& getParticleTypes() const { return types_; }
    const std::vector<int>&        getParticleTypeIdOfAllParticles() const { return typeIds_; }
    const std::vector<real>&       getCharges() const { return charges_; }
    const ExclusionLists&          getGmxExclusions() const { return exclusions_; }
    const NonBondedInteractionMap& getNonBondedInteractionMap() const { return interactions_; }
    CombinationRule                getCombinationRule() const { return rule_; }

    //! topology.cpp:264-268 (ParticleSequencer): the global index of particle
    //! particleName in residue residueName of the moleculeNr-th molecule
    //! called moleculeName (topologies from TopologyBuilder only)
    int sequenceID(const MoleculeName& moleculeName, int moleculeNr, const ResidueName& residueName,
                   const ParticleName& particleName) const;

private:
    friend class TopologyBuilder;
    std::vector<ParticleType> types_;
    std::vector<int>          typeIds_;
    std::vector<real>         charges_;
    ExclusionLists            exclusions_;
    NonBondedInteractionMap   interactions_;
    CombinationRule           rule_;
    std::map<std::tuple<std::string, int, std::string, std::string>, int> sequence_;
};

//! molecules.h:70-76
struct ParticleData
{
    std::string particleName_;
    std::string residueName_;
    std::string particleTypeName_;
    real        charge_;
};

//! molecules.h:80-170: a molecule type, its particles and the exclusions
//! between them (each particle excludes itself)
class Molecule
{
public:
    explicit Molecule(MoleculeName moleculeName);

    Molecule& addParticle(const ParticleName& particleName, const ResidueName& residueName, const Charge& charge,
                          const ParticleType& particleType);
    //! charge 0
    Molecule& addParticle(const ParticleName& particleName, const ResidueName& residueName,
                          const ParticleType& particleType);
    //! residue named after the molecule
    Molecule& addParticle(const ParticleName& particleName, const Charge& charge, const ParticleType& particleType);
    //! residue named after the molecule, charge 0
    Molecule& addParticle(const ParticleName& particleName, const ParticleType& particleType);

    //! both orders; a particle with itself is already excluded
    void addExclusion(int particleIndex, int particleIndexToExclude);
    //! by (particle name, residue name), resolved when the topology is built
    void addExclusion(std::tuple<std::string, std::string> particle,
                      std::tuple<std::string, std::string> particleToExclude);
    //! by particle name within the residue named after the molecule
    void addExclusion(const std::string& particleName, const std::string& particleNameToExclude);

    int                 numParticlesInMolecule() const { return static_cast<int>(particles_.size()); }
    const ParticleType& at(const std::string& particleTypeName) const;
    //! sorted, without duplicates, both orders, self pairs included
    std::vector<std::tuple<int, int>>             getExclusions() const;
    ParticleName                                  particleName(int i) const;
    ResidueName                                   residueName(int i) const;
    std::vector<ParticleData>                     particleData() const { return particles_; }
    std::unordered_map<std::string, ParticleType> particleTypesMap() const;
    MoleculeName                                  name() const { return name_; }

private:
    MoleculeName                      name_;
    std::vector<ParticleData>         particles_;
    std::vector<ParticleType>         particleTypes_; // distinct, in order of first use
    std::vector<std::tuple<int, int>> exclusions_;
    std::vector<std::tuple<std::string, std::string, std::string, std::string>> exclusionsByName_;
};

//! topology.h:133-175, topology.cpp:58-217
class TopologyBuilder
{
public:
    TopologyBuilder() = default;
    //! nMolecules copies of molecule, appended in order
    TopologyBuilder& addMolecule(const Molecule& molecule, int nMolecules);
    void             addParticleTypesInteractions(const ParticleTypesInteractions& particleTypesInteractions);
    //! the topology of every molecule added; throws if a type pair has no parameters
    Topology buildTopology();

private:
    std::vector<std::tuple<Molecule, int>> molecules_;
    std::vector<ParticleType>              particleTypes_; // distinct, in order of first appearance
    ParticleTypesInteractions              interactions_;
    int                                    numParticles_ = 0;
};

//! simulationstate.h:66-97
class SimulationState
{
public:
    SimulationState(const std::vector<Vec3>& coordinates, const std::vector<Vec3>& velocities,
                    const std::vector<Vec3>& forces, Box box, Topology topology);
    const Topology&          topology() const { return topology_; }
    Box                      box() const { return box_; }
    std::vector<Vec3>&       coordinates() { return x_; }
    const std::vector<Vec3>& coordinates() const { return x_; }
    std::vector<Vec3>&       velocities() { return v_; }
    std::vector<Vec3>&       forces() { return f_; }

private:
    std::vector<Vec3> x_, v_, f_;
    Box               box_;
    Topology          topology_;
};

//! forcecalculator.h:60-106
class ForceCalculator
{
public:
    ForceCalculator(const SimulationState& system, const NBKernelOptions& options);
    ~ForceCalculator();
    ForceCalculator(ForceCalculator&&) noexcept;
    ForceCalculator& operator=(ForceCalculator&&) noexcept;

    //! zeroes, then overwrites forces (gmxcalculator.cpp:70-83)
    void compute(ArrayRef<const Vec3> coordinates, ArrayRef<Vec3> forces);

    //! re-grid and rebuild the pair list (gmxcalculator.cpp:85-103)
    void updatePairList(ArrayRef<const int> particleInfoAllVdW, ArrayRef<Vec3> coordinates, const Box& box);

    //! {LJ, Coulomb} of the last compute when computeVirialAndEnergy is set (extension)
    std::array<double, 2> energies() const;
    //! virial of the last compute when computeVirialAndEnergy is set (extension)
    std::array<double, 9> virial() const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

//! integrator.h: leap-frog with put_atoms_in_box
class LeapFrog
{
public:
    LeapFrog(const Topology& topology, const Box& box);
    void integrate(real dt, ArrayRef<Vec3> x, ArrayRef<Vec3> v, ArrayRef<const Vec3> f);

private:
    std::vector<real> inverseMasses_;
    Box               box_;
};

} // namespace nblib
