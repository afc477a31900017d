// nblib-compatible host API over the HIP nonbonded engine; see nblib.h.
#include "nblib.h"

#include <algorithm>
#include <cmath>
#include <set>
#include <utility>

#include "nbx.h"

namespace nblib
{

namespace
{

// nbx status -> the exception nblib would throw for the same condition
void check(int rc, const char* what)
{
    if (rc == NBX_SUCCESS)
    {
        return;
    }
    const std::string msg = std::string(what) + ": " + nbx_last_error();
    if (rc == NBX_ERROR_INPUT || rc == NBX_ERROR_STATE)
    {
        throw InputException(msg);
    }
    throw std::runtime_error(msg);
}

bool finite(const std::vector<Vec3>& v)
{
    for (const Vec3& a : v)
    {
        if (!std::isfinite(a[0]) || !std::isfinite(a[1]) || !std::isfinite(a[2]))
        {
            return false;
        }
    }
    return true;
}

} // namespace

Box::Box(real l) : Box(l, l, l) {}

// box.cpp:52-63
Box::Box(real x, real y, real z)
{
    if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z))
    {
        throw InputException("Cannot have NaN or Inf box length.");
    }
    m_[0][0] = x;
    m_[1][1] = y;
    m_[2][2] = z;
}

void NonBondedInteractionMap::setInteractions(const std::string& a, const std::string& b, real c6, real c12)
{
    map_[{ a, b }] = { c6, c12 };
    map_[{ b, a }] = { c6, c12 };
}

real NonBondedInteractionMap::getC6(const std::string& a, const std::string& b) const
{
    auto it = map_.find({ a, b });
    if (it == map_.end())
    {
        throw InputException("No non-bonded interaction between " + a + " and " + b);
    }
    return it->second.first;
}

real NonBondedInteractionMap::getC12(const std::string& a, const std::string& b) const
{
    auto it = map_.find({ a, b });
    if (it == map_.end())
    {
        throw InputException("No non-bonded interaction between " + a + " and " + b);
    }
    return it->second.second;
}

// interactions.cpp:83-93: only the geometric rule is defined
static real combine(real v, real w, CombinationRule rule)
{
    if (rule != CombinationRule::Geometric)
    {
        throw InputException("unknown LJ Combination rule specified\n");
    }
    return std::sqrt(v * w);
}

// interactions.cpp:97-140
ParticleTypesInteractions& ParticleTypesInteractions::add(const std::string& name, real c6, real c12)
{
    auto ins = single_.insert({ name, { c6, c12 } });
    if (!ins.second && (ins.first->second.first != c6 || ins.first->second.second != c12))
    {
        throw InputException("Attempting to add nonbonded interaction parameters for particle type " + name
                             + " twice");
    }
    return *this;
}

ParticleTypesInteractions& ParticleTypesInteractions::add(const std::string& a, const std::string& b, real c6, real c12)
{
    auto ins = pairs_.insert({ { a, b }, { c6, c12 } });
    pairs_.insert({ { b, a }, { c6, c12 } });
    if (!ins.second && (ins.first->second.first != c6 || ins.first->second.second != c12))
    {
        throw InputException("Attempting to add nonbonded interaction parameters between the particle types "
                             + a + " " + b + " twice");
    }
    return *this;
}

// interactions.cpp:142-190
NonBondedInteractionMap ParticleTypesInteractions::generateTable() const
{
    NonBondedInteractionMap m;
    for (const auto& p1 : single_)
    {
        for (const auto& p2 : single_)
        {
            m.setInteractions(p1.first, p2.first, combine(p1.second.first, p2.second.first, rule_),
                              combine(p1.second.second, p2.second.second, rule_));
        }
    }
    for (const auto& p : pairs_)
    {
        m.setInteractions(p.first.first, p.first.second, p.second.first, p.second.second);
    }
    return m;
}

Topology::Topology(std::vector<ParticleType> particleTypes, std::vector<int> typeIdOfAllParticles,
                   std::vector<real> charges, ExclusionLists exclusions, NonBondedInteractionMap interactions,
                   CombinationRule rule) :
    types_(std::move(particleTypes)),
    typeIds_(std::move(typeIdOfAllParticles)),
    charges_(std::move(charges)),
    exclusions_(std::move(exclusions)),
    interactions_(std::move(interactions)),
    rule_(rule)
{
    const int n = numParticles();
    if (static_cast<int>(charges_.size()) != n || exclusions_.numLists() != n)
    {
        throw InputException("Topology: types, charges and exclusion lists must have one entry per particle");
    }
    for (int t : typeIds_)
    {
        if (t < 0 || t >= static_cast<int>(types_.size()))
        {
            throw InputException("Topology: particle type id out of range");
        }
    }
    // every type pair must have parameters (TopologyBuilder generates them)
    for (const ParticleType& a : types_)
    {
        for (const ParticleType& b : types_)
        {
            interactions_.getC6(a.name(), b.name());
        }
    }
}

// interactions.cpp:192-220
void ParticleTypesInteractions::merge(const ParticleTypesInteractions& other)
{
    for (const auto& p : other.single_)
    {
        add(p.first, p.second.first, p.second.second);
    }
    for (const auto& p : other.pairs_)
    {
        add(p.first.first, p.first.second, p.second.first, p.second.second);
    }
}

int Topology::sequenceID(const MoleculeName& moleculeName, int moleculeNr, const ResidueName& residueName,
                         const ParticleName& particleName) const
{
    auto it = sequence_.find({ moleculeName.value(), moleculeNr, residueName.value(), particleName.value() });
    if (it == sequence_.end())
    {
        throw InputException("no particle " + particleName.value() + " in residue " + residueName.value()
                             + " of molecule " + moleculeName.value() + " number " + std::to_string(moleculeNr));
    }
    return it->second;
}

// molecules.cpp:57-170
Molecule::Molecule(MoleculeName moleculeName) : name_(std::move(moleculeName)) {}

Molecule& Molecule::addParticle(const ParticleName& particleName, const ResidueName& residueName,
                                const Charge& charge, const ParticleType& particleType)
{
    auto same = [&](const ParticleType& t) { return t.name() == particleType.name(); };
    auto it   = std::find_if(particleTypes_.begin(), particleTypes_.end(), same);
    if (it == particleTypes_.end())
    {
        particleTypes_.push_back(particleType);
    }
    else if (!(*it == particleType))
    {
        throw InputException("Differing ParticleTypes with identical names encountered in the same molecule.");
    }
    particles_.push_back(ParticleData{ particleName.value(), residueName.value(), particleType.name(), charge.value() });
    // every particle excludes itself
    const int id = static_cast<int>(particles_.size()) - 1;
    exclusions_.emplace_back(id, id);
    return *this;
}

Molecule& Molecule::addParticle(const ParticleName& particleName, const ResidueName& residueName,
                                const ParticleType& particleType)
{
    return addParticle(particleName, residueName, Charge(0.0f), particleType);
}

Molecule& Molecule::addParticle(const ParticleName& particleName, const Charge& charge, const ParticleType& particleType)
{
    return addParticle(particleName, ResidueName(name_.value()), charge, particleType);
}

Molecule& Molecule::addParticle(const ParticleName& particleName, const ParticleType& particleType)
{
    return addParticle(particleName, ResidueName(name_.value()), Charge(0.0f), particleType);
}

void Molecule::addExclusion(int particleIndex, int particleIndexToExclude)
{
    const int n = numParticlesInMolecule();
    if (particleIndex < 0 || particleIndex >= n || particleIndexToExclude < 0 || particleIndexToExclude >= n)
    {
        throw InputException("exclusion between particle indices outside the molecule");
    }
    if (particleIndex != particleIndexToExclude)
    {
        exclusions_.emplace_back(particleIndex, particleIndexToExclude);
        exclusions_.emplace_back(particleIndexToExclude, particleIndex);
    }
}

void Molecule::addExclusion(std::tuple<std::string, std::string> particle,
                            std::tuple<std::string, std::string> particleToExclude)
{
    exclusionsByName_.emplace_back(std::get<0>(particle), std::get<1>(particle), std::get<0>(particleToExclude),
                                   std::get<1>(particleToExclude));
}

void Molecule::addExclusion(const std::string& particleName, const std::string& particleNameToExclude)
{
    addExclusion(std::make_tuple(particleName, name_.value()), std::make_tuple(particleNameToExclude, name_.value()));
}

const ParticleType& Molecule::at(const std::string& particleTypeName) const
{
    for (const ParticleType& t : particleTypes_)
    {
        if (t.name() == particleTypeName)
        {
            return t;
        }
    }
    throw std::out_of_range("no particle type " + particleTypeName + " in molecule " + name_.value());
}

ParticleName Molecule::particleName(int i) const
{
    return ParticleName(particles_.at(i).particleName_);
}

ResidueName Molecule::residueName(int i) const
{
    return ResidueName(particles_.at(i).residueName_);
}

std::unordered_map<std::string, ParticleType> Molecule::particleTypesMap() const
{
    std::unordered_map<std::string, ParticleType> m;
    for (const ParticleType& t : particleTypes_)
    {
        m.emplace(t.name(), t);
    }
    return m;
}

std::vector<std::tuple<int, int>> Molecule::getExclusions() const
{
    // (particle name, residue name) -> index; the names of a molecule's
    // particles identify them
    std::map<std::pair<std::string, std::string>, int> index;
    for (int i = 0; i < numParticlesInMolecule(); i++)
    {
        index.emplace(std::make_pair(particles_[i].particleName_, particles_[i].residueName_), i);
    }
    auto lookup = [&](const std::string& particle, const std::string& residue) {
        auto it = index.find({ particle, residue });
        if (it == index.end())
        {
            throw InputException("Particle " + particle + " in residue " + residue
                                 + " not found in list of particles");
        }
        return it->second;
    };
    std::set<std::tuple<int, int>> all(exclusions_.begin(), exclusions_.end());
    for (const auto& t : exclusionsByName_)
    {
        const int a = lookup(std::get<0>(t), std::get<1>(t));
        const int b = lookup(std::get<2>(t), std::get<3>(t));
        all.emplace(a, b);
        all.emplace(b, a);
    }
    return std::vector<std::tuple<int, int>>(all.begin(), all.end());
}

// topology.cpp:175-204
TopologyBuilder& TopologyBuilder::addMolecule(const Molecule& molecule, int nMolecules)
{
    if (nMolecules < 0)
    {
        throw InputException("cannot add a negative number of molecules");
    }
    for (const ParticleData& p : molecule.particleData())
    {
        const ParticleType& t    = molecule.at(p.particleTypeName_);
        auto                same = [&](const ParticleType& u) { return u.name() == t.name(); };
        auto                it   = std::find_if(particleTypes_.begin(), particleTypes_.end(), same);
        if (it == particleTypes_.end())
        {
            particleTypes_.push_back(t);
        }
        else if (!(*it == t))
        {
            throw InputException("Differing ParticleTypes with identical names encountered");
        }
    }
    molecules_.emplace_back(molecule, nMolecules);
    numParticles_ += nMolecules * molecule.numParticlesInMolecule();
    return *this;
}

void TopologyBuilder::addParticleTypesInteractions(const ParticleTypesInteractions& particleTypesInteractions)
{
    interactions_.merge(particleTypesInteractions);
}

// topology.cpp:58-174: per-particle charges and type ids, the exclusion
// lists of every copy of every molecule offset to global indices, the
// interaction table, and the check that every type pair has parameters
Topology TopologyBuilder::buildTopology()
{
    std::vector<int>  typeIds;
    std::vector<real> charges;
    typeIds.reserve(numParticles_);
    charges.reserve(numParticles_);
    std::map<std::string, int> typeId;
    for (size_t t = 0; t < particleTypes_.size(); t++)
    {
        typeId[particleTypes_[t].name()] = static_cast<int>(t);
    }
    ExclusionLists                                                         ex;
    std::map<std::tuple<std::string, int, std::string, std::string>, int> sequence;
    std::map<std::string, int>                                             copiesSoFar;
    int                                                                    offset = 0;
    for (const auto& mn : molecules_)
    {
        const Molecule&                         mol   = std::get<0>(mn);
        const int                               count = std::get<1>(mn);
        const int                               n     = mol.numParticlesInMolecule();
        const std::vector<std::tuple<int, int>> pairs = mol.getExclusions();
        const std::vector<ParticleData>         data  = mol.particleData();
        // per particle of the molecule: its excluded partners, sorted
        std::vector<std::vector<int>> perParticle(n);
        for (const auto& pr : pairs)
        {
            perParticle[std::get<0>(pr)].push_back(std::get<1>(pr));
        }
        int& copy = copiesSoFar[mol.name().value()];
        for (int m = 0; m < count; m++, copy++)
        {
            for (int i = 0; i < n; i++)
            {
                typeIds.push_back(typeId.at(data[i].particleTypeName_));
                charges.push_back(data[i].charge_);
                for (int j : perParticle[i])
                {
                    ex.indices.push_back(offset + j);
                }
                ex.offsets.push_back(static_cast<int>(ex.indices.size()));
                sequence[{ mol.name().value(), copy, data[i].residueName_, data[i].particleName_ }] = offset + i;
            }
            offset += n;
        }
    }
    NonBondedInteractionMap table = interactions_.generateTable();
    for (const ParticleType& a : particleTypes_)
    {
        for (const ParticleType& b : particleTypes_)
        {
            try
            {
                table.getC6(a.name(), b.name());
            }
            catch (const InputException&)
            {
                throw InputException("Missing nonbonded interaction parameters for pair " + a.name() + " "
                                     + b.name());
            }
        }
    }
    Topology top(particleTypes_, std::move(typeIds), std::move(charges), std::move(ex), std::move(table),
                 interactions_.getCombinationRule());
    top.sequence_ = std::move(sequence);
    return top;
}

// simulationstate.cpp:70-84
SimulationState::SimulationState(const std::vector<Vec3>& coordinates, const std::vector<Vec3>& velocities,
                                 const std::vector<Vec3>& forces, Box box, Topology topology) :
    x_(coordinates), v_(velocities), f_(forces), box_(box), topology_(std::move(topology))
{
    if (!finite(x_))
    {
        throw InputException("Input coordinates has at least one NaN");
    }
    if (!finite(v_))
    {
        throw InputException("Input velocities has at least one NaN");
    }
}

struct ForceCalculator::Impl
{
    nbx_context* ctx   = nullptr;
    unsigned     flags = NBX_COMPUTE_FORCES;
    int          n     = 0;
    ~Impl()
    {
        if (ctx)
        {
            nbx_destroy(ctx);
        }
    }
};

// GmxSetupDirector::setupGmxForceCalculator (gmxsetup.cpp:306-325)
ForceCalculator::ForceCalculator(const SimulationState& system, const NBKernelOptions& options) :
    impl_(std::make_unique<Impl>())
{
    const Topology& top = system.topology();
    nbx_options     o;
    nbx_default_options(&o);
    switch (options.coulombType)
    {
        case CoulombType::Pme: o.coulomb_type = NBX_COULOMB_EWALD; break;
        case CoulombType::Cutoff: o.coulomb_type = NBX_COULOMB_CUTOFF; break;
        case CoulombType::ReactionField: o.coulomb_type = NBX_COULOMB_REACTION_FIELD; break;
        default: throw InputException("Unsupported electrostatic interaction");
    }
    // the plain-C kernel always tabulates the Ewald correction
    // (gmxsetup.cpp:117-128); the others follow useTabulatedEwaldCorr
    const bool table   = options.nbnxmSimd == SimdKernels::SimdNo || options.useTabulatedEwaldCorr;
    o.ewald_correction = table ? NBX_EWALD_TABLE : NBX_EWALD_ANALYTICAL;
    o.pairlist_cutoff  = options.pairlistCutoff;
    o.device           = options.device;
    check(nbx_create(&o, &impl_->ctx), "ForceCalculator");

    const auto& types = top.getParticleTypes();
    const int   nt    = static_cast<int>(types.size());
    const auto& map   = top.getNonBondedInteractionMap();
    std::vector<double> c6(nt * nt), c12(nt * nt);
    for (int a = 0; a < nt; a++)
    {
        for (int b = 0; b < nt; b++)
        {
            c6[a * nt + b]  = map.getC6(types[a].name(), types[b].name());
            c12[a * nt + b] = map.getC12(types[a].name(), types[b].name());
        }
    }
    check(nbx_set_nonbonded_parameters(impl_->ctx, nt, c6.data(), c12.data()), "setNonBondedParameters");
    const ExclusionLists& ex = top.getGmxExclusions();
    impl_->n                 = top.numParticles();
    check(nbx_set_topology(impl_->ctx, impl_->n, top.getParticleTypeIdOfAllParticles().data(),
                           top.getCharges().data(), ex.offsets.data(), ex.indices.data()),
          "setAtomProperties");
    if (options.computeVirialAndEnergy)
    {
        impl_->flags |= NBX_COMPUTE_ENERGY | NBX_COMPUTE_VIRIAL;
    }
    const std::vector<Vec3>& x = system.coordinates();
    std::vector<int>         info(x.size(), 0);
    updatePairList(ArrayRef<const int>(info.data(), info.size()),
                   ArrayRef<Vec3>(const_cast<Vec3*>(x.data()), x.size()), system.box());
}

ForceCalculator::~ForceCalculator()                                = default;
ForceCalculator::ForceCalculator(ForceCalculator&&) noexcept            = default;
ForceCalculator& ForceCalculator::operator=(ForceCalculator&&) noexcept = default;

void ForceCalculator::compute(ArrayRef<const Vec3> coordinates, ArrayRef<Vec3> forces)
{
    if (static_cast<int>(coordinates.size()) != impl_->n || static_cast<int>(forces.size()) != impl_->n)
    {
        throw InputException("compute: coordinate and force arrays must hold one entry per particle");
    }
    static_assert(sizeof(Vec3) == 3 * sizeof(float), "Vec3 must be a packed float triplet");
    check(nbx_compute(impl_->ctx, &coordinates.data()->v[0], &forces.data()->v[0], impl_->flags, 0), "compute");
}

// gmxcalculator.cpp:85-103 (rectangular boxes only) and constructPairlist
void ForceCalculator::updatePairList(ArrayRef<const int> particleInfoAllVdW, ArrayRef<Vec3> coordinates,
                                     const Box& box)
{
    (void)particleInfoAllVdW; // every particle has LJ and charge on this path (gmxsetup.cpp:132-141)
    const auto& m = box.legacyMatrix();
    if (m[0][1] != 0 || m[0][2] != 0 || m[1][0] != 0 || m[1][2] != 0 || m[2][0] != 0 || m[2][1] != 0)
    {
        throw InputException("Only rectangular unit-cells are supported here");
    }
    if (static_cast<int>(coordinates.size()) != impl_->n)
    {
        throw InputException("updatePairList: one coordinate per particle expected");
    }
    const float diag[3] = { m[0][0], m[1][1], m[2][2] };
    check(nbx_put_on_grid(impl_->ctx, diag, &coordinates.data()->v[0], 0), "setParticlesOnGrid");
    check(nbx_construct_pairlist(impl_->ctx), "constructPairlist");
}

std::array<double, 2> ForceCalculator::energies() const
{
    std::array<double, 2> e{};
    check(nbx_get_energies(impl_->ctx, &e[0], &e[1]), "energies");
    return e;
}

std::array<double, 9> ForceCalculator::virial() const
{
    std::array<double, 9> v{};
    check(nbx_get_virial(impl_->ctx, v.data()), "virial");
    return v;
}

// integrator.cpp:53-75
LeapFrog::LeapFrog(const Topology& topology, const Box& box) : box_(box)
{
    inverseMasses_.resize(topology.numParticles());
    for (int i = 0; i < topology.numParticles(); i++)
    {
        const int t       = topology.getParticleTypeIdOfAllParticles()[i];
        inverseMasses_[i] = 1.0f / topology.getParticleTypes()[t].mass();
    }
}

void LeapFrog::integrate(real dt, ArrayRef<Vec3> x, ArrayRef<Vec3> v, ArrayRef<const Vec3> f)
{
    const auto& m = box_.legacyMatrix();
    for (std::size_t i = 0; i < x.size(); i++)
    {
        for (int d = 0; d < 3; d++)
        {
            v[i][d] += f[i][d] * dt * inverseMasses_[i];
            x[i][d] += v[i][d] * dt;
            // put_atoms_in_box for a rectangular box
            const real l = m[d][d];
            x[i][d] -= l * std::floor(x[i][d] / l);
            if (x[i][d] >= l)
            {
                x[i][d] -= l;
            }
        }
    }
}

} // namespace nblib
