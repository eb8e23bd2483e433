# Input set of the reference's RT diffraction test (test/diffraction/RT/input.jl),
# in the Python form that filament_amd.driver.prepare reads.
prefix = "results/"   # the path and prefix for the output files

file_initial_condition = "gauss.py"   # file with initial conditions
file_medium = "air.py"   # file with medium parameters

# Units
ru = 1e-3   # [m] unit of space in r direction
zu = 1.0   # [m] unit of space in z direction
tu = 1e-15   # [s] unit of time
Iu = 1e12 * 1e4   # [W/m**2] unit of intensity
rhou = 2.5e25   # [1/m**3] unit of plasma density

# Grid
geometry = "RT"
rmax = 10.0   # [ru]
Nr = 500
tmin, tmax = -400.0, 400.0   # [tu]
Nt = 4096

# Model
zmax = 10.0   # [zu]
dz_initial = zmax / 100.0
NONLINEARITY = False
PLASMA = False
KPARAXIAL = False
QPARAXIAL = True

rguard = 1.0   # [ru]
tguard = 20.0   # [tu]
kguard = 45.0   # [degrees]
wguard = 1e16   # [1/s]

dzphimax = pi / 100.0
Istop = 1e3   # [Iu]

HTLOAD = False
file_ht = "plan_dht.jld2"
ALG = RK3()

dz_plothdf = zmax
z = 0.0   # [m]
