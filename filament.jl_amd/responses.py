"""The built-in nonlinear responses and the plasma equation, named as in the
reference's medium files (``"init" => init_cubic`` etc.,
test/diffraction/RT/air.jl:36-108).  User Julia closures cannot cross the
C ABI, so a medium file selects one of these kinds; the library evaluates the
matching ``calculate`` on the GPU."""


class _Kind:
    __slots__ = ("name",)

    def __init__(self, name):
        self.name = name

    def __repr__(self):
        return f"<response {self.name}>"


init_cubic = _Kind("cubic")                     # medium_responses/cubic.jl
init_current_free = _Kind("current_free")       # medium_responses/current_free.jl
init_current_losses = _Kind("current_losses")   # medium_responses/current_losses.jl
init_raman = _Kind("raman")                     # medium_responses/raman.jl
init_photoionization_avalanche = _Kind("photoionization_avalanche")   # plasma_equations/


def response_kind(init):
    if isinstance(init, _Kind):
        return init.name
    if isinstance(init, str):
        return init
    raise TypeError(f"unsupported response init {init!r}")


class RK3:
    """Placeholder for ODEIntegrators.RK3() in input files: Kutta's
    third-order method, the tableau of every geometry."""


class RK4:
    """Placeholder for ODEIntegrators.RK4() in input files: the classic
    fourth-order method, integrated for the spatial geometries R and XY (the
    reference's R and XY examples select it, examples/R/input.jl:38)."""


def alg_order(alg):
    """3 for RK3() (or no ALG), 4 for RK4(); other integrators are not built."""
    if alg is None or isinstance(alg, RK3) or alg is RK3:
        return 3
    if isinstance(alg, RK4) or alg is RK4:
        return 4
    raise NotImplementedError(f"ALG {alg!r}: the library integrates RK3 and (R, XY) RK4")
