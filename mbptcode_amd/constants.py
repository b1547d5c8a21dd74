"""Named numbers of the ISDF BSE sigma-vector / space-time chi0 path.

Each value is the one the reference reads (src/Base/numerics/constants.py, line
given per entry), so a caller moving between the two sees the same defaults.
"""

#: 1 GB as the tile budgets count it (constants.py:17).
BYTES_PER_GB = 1e9

#: Hartree-term prefactor per spin channel: singlet 2, triplet 0 (constants.py:287).
KAPPA = {'singlet': 2.0, 'triplet': 0.0}

#: Scratch budget of the block action, in GB (constants.py:292).  The
#: reference sizes its grid-row tiles by it; the GPU action applies trial
#: vectors in slices whose device workspace fits it.
ISDF_TILE_GB = 4.0

#: The BSE Davidson's roots and residual tolerance (constants.py:649-650), and
#: its cycle cap, three trial-space cycles of 50 (constants.py:1340).
BSE_DAVIDSON_NROOTS = 5
BSE_DAVIDSON_CONV_TOL = 1e-8
BSE_FORCE_MAX_CYCLE = 150

#: The Casida eigensolver choices (constants.py:1865), the root adjoint forms
#: (:1243) and the root-following modes (:1654) a `BSENumerics` accepts.
BSE_SOLVERS = ('auto', 'dense', 'davidson')
BSE_ADJOINTS = ('explicit', 'grid')
ROOT_FOLLOWING = ('overlap',)

#: Time points of the self-energy's contour grid (constants.py:493), Pade
#: points of the space-time self-energy (:1849) and the scale of their
#: Gauss-Legendre grid (:1851).
CD_NTAU = 24
DEFAULT_NPADE = 16
GL_FREQUENCY_SCALE = 1.0

#: Gauss-Legendre frequencies of the dRPA energy quadrature (constants.py:469).
RPA_ENERGY_NFREQ = 40

#: Truncated-SVD cut of the minimax transform fit (constants.py:1148).
TRANSFORM_FIT_RCOND = 1e-100

#: Laplace-test target the automatic tau-count resolver aims for (constants.py:1844).
DEFAULT_TAU_TARGET = 1e-10

# Davidson (constants.py:1170-2193)
DAVIDSON_FLOOR_EPS_MULTIPLE = 1e3       # :1170 residual floor, eps multiples
DAVIDSON_TOL_RESIDUAL = 1e-6            # :1194
DAVIDSON_MAX_CYCLE = 200                # :1195
DAVIDSON_LINDEP_FLOOR = 1e-22           # :1196
DAVIDSON_LINDEP_CEILING = 1e-14         # :1197
DAVIDSON_NULL_DIRECTION = 1e-8          # :1198
DAVIDSON_PRECONDITIONER = 'screened'    # :1380
DAVIDSON_PRECONDITIONERS = ('bare', 'screened', 'learned')  # :1381 + the MLP
DAVIDSON_EXTRA_ROOTS_FLOOR = 2          # :1648
DAVIDSON_EXTRA_ROOTS_DIVISOR = 2        # :1650
DAVIDSON_GUESS_FACTOR = 2               # :1886
DAVIDSON_GUESS_DEG_TOL = 1e-3           # :2190
DAVIDSON_DENOMINATOR_FLOOR = 1e-8       # :2193
DAVIDSON_COMPLETION_SEED = 0            # :2124 random directions of the completion solve
DAVIDSON_FOLLOW_OVERLAP_TIE = 1e-6      # :2129 overlaps this close are a tie

#: Smallest Rayleigh matrix whose Ritz pairs the device eigensolver
#: (ritz.hip) computes in place of host LAPACK.  Inside the Davidson the host
#: step also pays for its products yTGy, yTHy and the upload of the vectors;
#: on one MI355X the plain FP64 solve at 100/700/6000 (Rayleigh matrices of 40
#: to 161) took 64.3 ms with the host step throughout and 63.2 ms with the
#: device solver from 70 or from 40.  Not a reference constant.
DAVIDSON_DEVICE_RITZ_MIN = 64

#: Largest pair space a Davidson whose trial space ran out is completed in
#: by a dense solve (constants.py:590).
BSE_DENSE_MAX_NOV = 12000

# the ISDF fit (constants.py:36, :1755): the metric root's relative eigenvalue
# floor and the Tikhonov floor of the balanced Gram matrix
AUX_METRIC_ROOT_FLOOR = 1e-12
ISDF_FIT_REGULARIZATION = 4e-7

# the (A - B) probe (constants.py:253-254, :1234)
AMB_PROBE_TOL = 1e-6                    # residual bar of the converged probe, Ha
AMB_PROBE_SIGN_TOLS = (1e-2, 1e-3, 1e-4)  # relative bars of the sign passes
PROBE_START_MIX = 0.1                   # share of 1/d in the warm sign pass

#: Hartree -> meV and -> eV.
HARTREE_TO_MEV = 27211.386245988
HARTREE_TO_EV = 27.211386245988

# seconds git may take to name the commit a record was made on
# (Base/numerics/constants.py:1112)
GIT_PROVENANCE_TIMEOUT = 10.0
