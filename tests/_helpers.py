"""Test helpers: the same problem as an oracle configuration."""
import numpy as np


def load_pkg():
    """Import the filament.jl_amd package under the name filament_amd."""
    import importlib.util
    import os
    import sys
    if "filament_amd" in sys.modules:
        return sys.modules["filament_amd"]
    root = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "filament.jl_amd")
    spec = importlib.util.spec_from_file_location("filament_amd", os.path.join(root, "__init__.py"),
                                                  submodule_search_locations=[root])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["filament_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def load_oracle():
    import os
    import sys
    d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle")
    if d not in sys.path:
        sys.path.insert(0, d)
    import oracle
    return oracle


def oracle_config(p):
    """The same problem as an oracle configuration dict (tests only)."""
    resp = {response_name(r["init"]): r for r in p["responses"]} if p["NONLINEARITY"] else {}
    pe = p["plasma_equation"]
    comps = pe["components"]
    tabs = [tuple(np.asarray(v, dtype=float) for v in c["ionization_rate"]) for c in comps]
    medium = 0 if p["permittivity"] is load_pkg().physics.peck_reeder_air else None
    if medium is None:
        raise ValueError("the oracle restates only Peck-Reeder air")
    cub = resp.get("cubic")
    ram = resp.get("raman")
    if ram is not None and ram["raman_response"] is not load_pkg().physics.mlejnek_raman_response:
        raise ValueError("the oracle restates only the Mlejnek Raman response of air")
    for r in resp.values():
        if "n2" in r and r["n2"] != p["n2"]:
            raise ValueError("the oracle takes one n2 for the cubic and Raman responses")
    return dict(
        geometry=0, Nr=p["Nr"], Nt=p["Nt"], rmax=p["rmax"], tmin=p["tmin"], tmax=p["tmax"],
        ru=p["ru"], zu=p["zu"], tu=p["tu"], Iu=p["Iu"], rhou=p["rhou"], lam0=p["lam0"],
        medium=medium, n_const=1.0, n2=p["n2"],
        nonlinearity=int(p["NONLINEARITY"]), plasma=int(p["PLASMA"]),
        kparaxial=int(p["KPARAXIAL"]), qparaxial=int(p["QPARAXIAL"]),
        rguard=p["rguard"], tguard=p["tguard"], kguard=p["kguard"], wguard=p["wguard"],
        cubic=(0 if cub is None else (1 if cub["THG"] else 2)), raman=(0 if ram is None else (1 if ram["THG"] else 2)),
        current_free=int("current_free" in resp),
        nuc=resp["current_free"]["nuc"] if "current_free" in resp else p["nuc"],
        mr=resp["current_free"]["mr"] if "current_free" in resp else p["mr"],
        current_losses=int("current_losses" in resp),
        losses_ereal=int(resp["current_losses"]["EREAL"]) if "current_losses" in resp else 0,
        pe_ereal=int(pe["EREAL"]), pe_kdep=int(pe["KDEP"]), pe_rho0=pe["rho0"], pe_rho_nt=pe["rho_nt"],
        pe_nuc=pe["nuc"], pe_mr=pe["mr"],
        comp_frac=[c["fraction"] for c in comps], comp_Ui=[c["ionization_energy"] for c in comps],
        tables=tabs,
    )


def oracle_config_t(p):
    """The same T problem as an oracle configuration dict (tests only)."""
    c = oracle_config(dict(p, rmax=0.0, Nr=0, ru=1.0, rguard=0.0, kguard=0.0))
    c.update(geometry=1, Nr=0)
    return c


def oracle_config_xyt(p, guard3d_bug=False):
    """The same XYT problem as an oracle configuration dict (tests only)."""
    c = oracle_config(dict(p, rmax=0.0, Nr=0, ru=p["xu"], rguard=0.0, kguard=p["kxguard"]))
    c.update(geometry=2, Nr=0, Nx=p["Nx"], Ny=p["Ny"], xmin=p["xmin"], xmax=p["xmax"], ymin=p["ymin"],
             ymax=p["ymax"], xu=p["xu"], yu=p["yu"], xguard=p["xguard"], yguard=p["yguard"],
             kxguard=p["kxguard"], kyguard=p["kyguard"], guard3d_bug=int(guard3d_bug))
    return c


def xyt_rows(E):
    """[Nx, Ny, Nt] -> [Nx*Ny, Nt] with x fastest (the oracle's and the device's row order)."""
    E = np.asarray(E)
    return E.reshape(E.shape[0] * E.shape[1], E.shape[2], order="F")


def response_name(init):
    return load_pkg().responses.response_kind(init)


def spatial_oracle(p, precision=8, E0=None):
    """The oracle of an R or XY parameter dict (filament_amd.workloads.spatial_air
    or driver.prepare output), with E0 the initial field on the grid (default:
    p's initial_condition on the oracle's own grid)."""
    o = load_oracle()
    pk = load_pkg()
    c = o.OrcSpatialConfig()
    xy = p["geometry"] == "XY"
    c.geometry = o.GEOM_XY if xy else o.GEOM_R
    if xy:
        c.Nx, c.Ny = p["Nx"], p["Ny"]
        c.xmin, c.xmax, c.ymin, c.ymax, c.xu, c.yu = p["xmin"], p["xmax"], p["ymin"], p["ymax"], p["xu"], p["yu"]
        c.xguard, c.yguard, c.kxguard, c.kyguard = p["xguard"], p["yguard"], p["kxguard"], p["kyguard"]
    else:
        c.Nr, c.rmax, c.ru = p["Nr"], p["rmax"], p["ru"]
        c.rguard, c.kguard = p["rguard"], p["kguard"]
    c.zu, c.Iu, c.lam0 = p["zu"], p["Iu"], p["lam0"]
    c.medium = o.MEDIUM_AIR
    c.n2 = p["n2"]
    c.nonlinearity = int(bool(p["NONLINEARITY"]))
    c.kparaxial, c.qparaxial = int(bool(p["KPARAXIAL"])), int(bool(p["QPARAXIAL"]))
    c.alg = pk.responses.alg_order(p.get("ALG"))
    if E0 is None:
        if xy:
            x = np.linspace(p["xmin"], p["xmax"], p["Nx"])
            y = np.linspace(p["ymin"], p["ymax"], p["Ny"])
            E0 = p["initial_condition"](x, y, p["xu"], p["yu"], p["Iu"])
        else:
            r = o.bessel_j0_zeros(p["Nr"] + 1)
            r = r[:-1] * p["rmax"] / r[-1]
            E0 = p["initial_condition"](r, p["ru"], p["Iu"])
    return o.SpatialOracle(c, E0, precision)
