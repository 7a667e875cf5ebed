"""MI355X-native randomized numerical linear algebra: the hot path of PARLA 0.1.4.

Drop-in names (same constructors and call signatures as ``parla``):
sketching operators (``SkOpGA``, ``SkOpSJ``, ``SkOpIN``, ``gaussian_operator``,
``sjlt_operator``, ...), sketch-and-precondition least squares (``SPO``,
``PcSS2``, ``lsqr``, ``a_lift_precond``) and the QB chain (``RS1``, ``RF1``,
``QB1``, ``SVD1``, ``EVD1``, ``orth``, ``lu_stabilize``).  Inputs may be numpy
arrays (results come back as numpy) or CUDA tensors (results stay on the
device).  All arithmetic runs in the HIP library ``libparla_amd.so``.
"""
__version__ = '0.1.4+mi355x'


def __getattr__(name):
    # lazy: importing the package must not require a GPU or the built library
    import importlib
    table = {
        'gaussian_operator': 'parla_amd.utils.sketching', 'sjlt_operator': 'parla_amd.utils.sketching',
        'orthonormal_operator': 'parla_amd.utils.sketching', 'sampling_operator': 'parla_amd.utils.sketching',
        'SketchOpGen': 'parla_amd.comps.sketchers.oblivious', 'SkOpGA': 'parla_amd.comps.sketchers.oblivious',
        'SkOpSJ': 'parla_amd.comps.sketchers.oblivious', 'SkOpIN': 'parla_amd.comps.sketchers.oblivious',
        'SkOpON': 'parla_amd.comps.sketchers.oblivious',
        'RS1': 'parla_amd.comps.sketchers.aware', 'RowSketcher': 'parla_amd.comps.sketchers.aware',
        'RF1': 'parla_amd.comps.rangefinders', 'RangeFinder': 'parla_amd.comps.rangefinders',
        'QB1': 'parla_amd.comps.qb', 'QBDecomposer': 'parla_amd.comps.qb',
        'SPO': 'parla_amd.drivers.least_squares', 'OverLstsqSolver': 'parla_amd.drivers.least_squares',
        'SPOMixed': 'parla_amd.drivers.mixed_ls', 'SplitOperator': 'parla_amd.utils.split_operator',
        'SketchAndPrecondLog': 'parla_amd.comps.determiter.logging',
        'SVD1': 'parla_amd.drivers.svd', 'SVDecomposer': 'parla_amd.drivers.svd',
        'EVD1': 'parla_amd.drivers.evd', 'EVDecomposer': 'parla_amd.drivers.evd',
        'PcSS2': 'parla_amd.comps.determiter.saddle', 'lsqr': 'parla_amd.comps.determiter.lsqr',
        'a_lift_precond': 'parla_amd.comps.preconditioning',
        'orth': 'parla_amd.utils.linalg_wrappers', 'lu_stabilize': 'parla_amd.utils.linalg_wrappers',
        'SkOpSS': 'parla_amd.comps.sketchers.oblivious', 'SkOpTC': 'parla_amd.comps.sketchers.oblivious',
        'srct_operator': 'parla_amd.utils.sketching', 'sparse_sign_operator': 'parla_amd.utils.sketching',
        'apply_srct': 'parla_amd.utils.sketching', 'generate_srct': 'parla_amd.utils.sketching',
        'SSO1': 'parla_amd.drivers.least_squares', 'sso1': 'parla_amd.drivers.least_squares',
        'spo1': 'parla_amd.drivers.least_squares', 'spo3': 'parla_amd.drivers.least_squares',
        'SPU1': 'parla_amd.drivers.least_squares', 'spu1': 'parla_amd.drivers.least_squares',
        'UnderLstsqSolver': 'parla_amd.drivers.least_squares',
        'svd_right_precond': 'parla_amd.comps.preconditioning', 'a_lift': 'parla_amd.comps.preconditioning',
        'PrecondSaddleSolver': 'parla_amd.comps.determiter.saddle',
        'QB2': 'parla_amd.comps.qb', 'QB3': 'parla_amd.comps.qb', 'qb': 'parla_amd.comps.qb',
        'EVD2': 'parla_amd.drivers.evd', 'evd1': 'parla_amd.drivers.evd', 'evd2': 'parla_amd.drivers.evd',
        'svd1': 'parla_amd.drivers.svd',
        'PcSS1': 'parla_amd.comps.determiter.saddle', 'pcg': 'parla_amd.comps.determiter.pcg',
        'SPS1': 'parla_amd.drivers.saddlesys', 'SPS2': 'parla_amd.drivers.saddlesys',
        'SaddleSolver': 'parla_amd.drivers.saddlesys', 'sps': 'parla_amd.drivers.saddlesys',
    }
    if name in table:
        return getattr(importlib.import_module(table[name]), name)
    raise AttributeError(name)
