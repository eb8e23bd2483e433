# Medium of the reference's RT tests (test/diffraction/RT/air.jl): Peck-Reeder
# air.  The tests run with NONLINEARITY = PLASMA = false, so the response and
# plasma definitions below are read but not used.


def permittivity(w):
    # E.R. Peck and K. Reeder "Dispersion of Air" JOSA, 62, 958 (1972)
    if w == 0.0:
        sig = 0.0
    else:
        lam = 2.0 * pi * C0 / w
        sig = 1.0 / lam
        sig = sig * 1e-6   # 1/m -> 1/um
    dum = 8060.51 + 2480990.0 / (132.247 - sig ** 2) + 17455.7 / (39.32957 - sig ** 2)
    n = 1.0 + dum * 1e-8
    return n ** 2


def permeability(w):
    return 1.0


n2 = 0.5e-23   # [m**2/W] nonlinear index
nuc = 5e12   # [1/s] collision frequency
mr = 1.0   # [me] reduced mass of electron and hole

cubic = {"init": init_cubic, "THG": True, "n2": n2}
current_free = {"init": init_current_free, "nuc": nuc, "mr": mr}
current_losses = {"init": init_current_losses, "EREAL": False}
responses = [cubic, current_free, current_losses]

N2 = {"fraction": 0.79, "ionization_energy": 15.576, "ionization_rate": "../reference/multiphoton_N2.tf"}
O2 = {"fraction": 0.21, "ionization_energy": 12.063, "ionization_rate": "../reference/multiphoton_O2.tf"}
components = [N2, O2]
plasma_equation = {
    "init": init_photoionization_avalanche, "ALG": RK3(), "EREAL": False, "KDEP": True,
    "rho0": 0.0, "rho_nt": 2.5e25, "nuc": nuc, "mr": mr, "components": components,
}
