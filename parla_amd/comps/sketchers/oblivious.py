"""Data-oblivious sketching operator generators (drop-in for
parla/comps/sketchers/oblivious.py).  ``gen(n_rows, n_cols, rng)`` returns a
device operator (see parla_amd.utils.sketching)."""
import parla_amd.utils.sketching as usk


class SketchOpGen:
    """oblivious.py:25-28."""

    def __call__(self, n_rows, n_cols, rng):
        raise NotImplementedError()


class SkOpON(SketchOpGen):
    """Orthonormal (Haar) sketching operator, oblivious.py:31-35."""

    def __call__(self, n_rows, n_cols, rng):
        return usk.orthonormal_operator(n_rows, n_cols, rng)


class SkOpGA(SketchOpGen):
    """Gaussian sketching operator, oblivious.py:38-45.

    ``backend`` selects where the values are drawn ('numpy': the reference's
    exact stream, 'device': Philox on the GPU); None uses the module default.
    """

    def __init__(self, normalize=True, backend=None):
        self.normalize = normalize
        self.backend = backend

    def __call__(self, n_rows, n_cols, rng):
        return usk.gaussian_operator(n_rows, n_cols, rng, self.normalize, backend=self.backend)


class SkOpSJ(SketchOpGen):
    """SJLT (sparse sign, vec_nnz per column), oblivious.py:48-55."""

    def __init__(self, vec_nnz=8, backend=None):
        self.vec_nnz = vec_nnz
        self.backend = backend

    def __call__(self, n_rows, n_cols, rng):
        return usk.sjlt_operator(n_rows, n_cols, rng, self.vec_nnz, backend=self.backend)


class SkOpSS(SketchOpGen):
    """Sparse sign operator, oblivious.py:58-65 (the reference's __call__
    drops its result; this one returns it)."""

    def __init__(self, density=0.05):
        self.density = density

    def __call__(self, n_rows, n_cols, rng):
        return usk.sparse_sign_operator(n_rows, n_cols, rng, self.density)


class SkOpTC(SketchOpGen):
    """SRTT based on the DCT-II (the SRCT), oblivious.py:68-72."""

    def __call__(self, n_rows, n_cols, rng):
        return usk.srct_operator(n_rows, n_cols, rng)


class SkOpIN(SketchOpGen):
    """Row / column selection, oblivious.py:75-82."""

    def __init__(self, indices=None):
        self.indices = indices

    def __call__(self, n_rows, n_cols, rng):
        return usk.sampling_operator(n_rows, n_cols, rng, self.indices)
