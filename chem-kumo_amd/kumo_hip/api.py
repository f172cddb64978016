"""Python mirror of the reference's public Evaluator vocabulary
(include/chem/kumo/evaluator.hpp:112-472, include/chem/kumo/spin.hpp).
The numbering of every enum is the reference's, so the same integers cross
the C-ABI (include/kumo_hip.h)."""
import enum


class DensityComponent(enum.IntEnum):  # include/chem/kumo/spin.hpp:8
    Total = 0
    Alpha = 1
    Beta = 2
    Difference = 3


class SpinChannel(enum.IntEnum):  # include/chem/kumo/spin.hpp:12
    Alpha = 0
    Beta = 1


class PropertyKind(enum.IntEnum):  # include/chem/kumo/evaluator.hpp:112-150
    Density = 0
    Gradient = 1
    Hessian = 2
    Laplacian = 3
    Tau = 4
    Orbital = 5
    Elf = 6
    ReducedGradient = 7
    SignLambda2Density = 8
    ElectrostaticPotential = 9
    Lol = 10
    InteractionRegionIndicator = 11
    DensityOverlapRegionsIndicator = 12
    HamiltonianKineticEnergy = 13
    PotentialEnergyDensity = 14
    EnergyDensity = 15
    SpinPolarization = 16
    ShannonEntropyDensity = 17
    FisherInformationDensity = 18
    AverageLocalIonizationEnergy = 19
    LocalElectronAffinity = 20
    LocalElectronAttachmentEnergy = 21
    LocalElectronegativity = 22
    LocalHardness = 23
    OrbitalGradient = 24
    OrbitalHessian = 25
    OrbitalWeightedFukuiPlus = 26
    OrbitalWeightedFukuiMinus = 27
    OrbitalWeightedFukuiZero = 28
    OrbitalWeightedDualDescriptor = 29
    TauComponents = 30
    HamiltonianKineticEnergyComponents = 31
    PotentialKineticRatio = 32
    EnergyDensityPerElectron = 33
    KineticEnergyPerElectron = 34
    ElfTsirelsonStash = 35
    LolTsirelsonStash = 36


class Backend(enum.IntEnum):
    """Backend {AUTO, AO_P, AO_MO} of include/chem/kumo/evaluator.hpp:403; the
    HIP build keeps these as the density-contraction choice and runs them on
    the GPU."""
    AUTO = 0
    AO_P = 1
    AO_MO = 2


class Precision(enum.IntEnum):
    """Arithmetic of the dense contractions on the GPU."""
    FP64 = 0
    FP32 = 1
    BF16X3 = 2   # two-term bf16 split, three MFMA products, FP32 accumulation
    BF16X6 = 3   # three-term bf16 split, six MFMA products, FP32 accumulation


def channels(kind):
    """Components of a property (src/evaluator.cpp:30-40)."""
    k = int(kind)
    if k in (PropertyKind.Gradient, PropertyKind.OrbitalGradient, PropertyKind.TauComponents,
             PropertyKind.HamiltonianKineticEnergyComponents):
        return 3
    if k in (PropertyKind.Hessian, PropertyKind.OrbitalHessian):
        return 6
    return 1


class Property(tuple):
    """(kind, component, channel, orbital_index) — the key of
    include/chem/kumo/evaluator.hpp:166-180."""

    def __new__(cls, kind, component=DensityComponent.Total, channel=SpinChannel.Alpha, orbital_index=0):
        return super().__new__(cls, (int(kind), int(component), int(channel), int(orbital_index)))

    kind = property(lambda s: s[0])
    component = property(lambda s: s[1])
    channel = property(lambda s: s[2])
    orbital_index = property(lambda s: s[3])


class _Density:
    """density::total.value() and friends (evaluator.hpp:183-318)."""

    def __init__(self, component):
        self.component = component

    def _p(self, kind):
        return Property(kind, self.component)

    def value(self): return self._p(PropertyKind.Density)
    def gradient(self): return self._p(PropertyKind.Gradient)
    def hessian(self): return self._p(PropertyKind.Hessian)
    def laplacian(self): return self._p(PropertyKind.Laplacian)
    def tau(self): return self._p(PropertyKind.Tau)
    def elf(self): return self._p(PropertyKind.Elf)
    def reduced_gradient(self): return self._p(PropertyKind.ReducedGradient)
    def sign_lambda2_density(self): return self._p(PropertyKind.SignLambda2Density)
    def electrostatic_potential(self): return self._p(PropertyKind.ElectrostaticPotential)
    def lol(self): return self._p(PropertyKind.Lol)
    def interaction_region_indicator(self): return self._p(PropertyKind.InteractionRegionIndicator)
    def density_overlap_regions_indicator(self): return self._p(PropertyKind.DensityOverlapRegionsIndicator)
    def hamiltonian_kinetic_energy(self): return self._p(PropertyKind.HamiltonianKineticEnergy)
    def potential_energy_density(self): return self._p(PropertyKind.PotentialEnergyDensity)
    def energy_density(self): return self._p(PropertyKind.EnergyDensity)
    def tau_components(self): return self._p(PropertyKind.TauComponents)
    def hamiltonian_kinetic_energy_components(self): return self._p(PropertyKind.HamiltonianKineticEnergyComponents)
    def potential_kinetic_ratio(self): return self._p(PropertyKind.PotentialKineticRatio)
    def energy_density_per_electron(self): return self._p(PropertyKind.EnergyDensityPerElectron)
    def elf_tsirelson_stash(self): return self._p(PropertyKind.ElfTsirelsonStash)
    def lol_tsirelson_stash(self): return self._p(PropertyKind.LolTsirelsonStash)
    def kinetic_energy_per_electron(self): return self._p(PropertyKind.KineticEnergyPerElectron)
    def spin_polarization(self): return self._p(PropertyKind.SpinPolarization)
    def shannon_entropy_density(self): return self._p(PropertyKind.ShannonEntropyDensity)
    def fisher_information_density(self): return self._p(PropertyKind.FisherInformationDensity)
    def average_local_ionization_energy(self): return self._p(PropertyKind.AverageLocalIonizationEnergy)
    def local_electron_affinity(self): return self._p(PropertyKind.LocalElectronAffinity)
    def local_electron_attachment_energy(self): return self._p(PropertyKind.LocalElectronAttachmentEnergy)
    def local_electronegativity(self): return self._p(PropertyKind.LocalElectronegativity)
    def local_hardness(self): return self._p(PropertyKind.LocalHardness)
    def orbital_weighted_fukui_plus(self): return self._p(PropertyKind.OrbitalWeightedFukuiPlus)
    def orbital_weighted_fukui_minus(self): return self._p(PropertyKind.OrbitalWeightedFukuiMinus)
    def orbital_weighted_fukui_zero(self): return self._p(PropertyKind.OrbitalWeightedFukuiZero)
    def orbital_weighted_dual_descriptor(self): return self._p(PropertyKind.OrbitalWeightedDualDescriptor)


class density:  # namespace density (evaluator.hpp:321-327)
    total = _Density(DensityComponent.Total)
    alpha = _Density(DensityComponent.Alpha)
    beta = _Density(DensityComponent.Beta)
    difference = _Density(DensityComponent.Difference)


class _Orbital:
    def __init__(self, channel, index):
        self.channel, self.index = channel, index

    def value(self): return Property(PropertyKind.Orbital, DensityComponent.Total, self.channel, self.index)
    def gradient(self): return Property(PropertyKind.OrbitalGradient, DensityComponent.Total, self.channel, self.index)
    def hessian(self): return Property(PropertyKind.OrbitalHessian, DensityComponent.Total, self.channel, self.index)


class orbital:  # namespace orbital (evaluator.hpp:349-357)
    @staticmethod
    def alpha(index): return _Orbital(SpinChannel.Alpha, index)

    @staticmethod
    def beta(index): return _Orbital(SpinChannel.Beta, index)


class Request:
    """A list of properties (evaluator.hpp:361-374)."""

    def __init__(self, *props):
        self.properties = list(props)

    def add(self, prop):
        self.properties.append(Property(*prop))
        return self
