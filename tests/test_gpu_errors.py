"""Argument errors of the C ABI and the Solver on a GPU: every case is caught on
the host before any kernel runs, reported through spr_last_error (SpringHipError
in Python, exceptions in the Solver), and leaves the handle usable."""
import numpy as np
import pytest

from helpers import case

pytestmark = pytest.mark.gpu


def test_capi_argument_errors(hip):
    amp, rho, supp = case(64, 3, seed=41)
    with pytest.raises(hip.SpringHipError, match="unsupported grid size"):
        hip.Backend(np.zeros((96, 96), np.float32), capacity=1)
    be = hip.Backend(amp, capacity=3)
    with pytest.raises(hip.SpringHipError, match="capacity"):
        be.set_population(np.concatenate([rho, rho[:1]]), np.concatenate([supp, supp[:1]]))
    be.set_population(rho, supp)
    with pytest.raises(hip.SpringHipError, match="negative iteration"):
        be.iterate("HIO", -1, 0.9, 0.0)
    with pytest.raises(hip.SpringHipError, match="unknown algorithm"):
        be.iterate(7, 1, 0.9, 0.0)
    with pytest.raises(hip.SpringHipError, match="slot index out of range"):
        be.gather([0, 5, 1])
    with pytest.raises(hip.SpringHipError, match="slot index out of range"):
        be.fetch_individuals([3])
    with pytest.raises(hip.SpringHipError, match="more individuals requested"):
        be.get_population(P=4)
    with pytest.raises(hip.SpringHipError, match="negative timing stride"):
        be.timing(True, stride=-2)
    # the handle still works
    be.iterate("ER", 2, 0.0, 0.0)
    got = be.get_population(support=False)
    assert np.all(np.isfinite(got))


def test_empty_population_is_a_no_op(hip):
    amp, rho, supp = case(64, 1, seed=42)
    be = hip.Backend(amp, capacity=2)
    be.set_population(rho[:0], supp[:0])
    be.iterate("HIO", 3, 0.9, 0.0)
    be.sw(2.0, 0.1)
    assert be.evaluate()["error_fourier"].shape == (0,)


def test_solver_argument_errors():
    from s_spi_spring_amd import spring_core
    pr = spring_core.phase_retrieval
    I = np.ones((64, 64), np.float32)
    with pytest.raises(Exception, match="unsupported pattern size"):
        pr.Solver(np.ones((96, 96), np.float32), gpus=[0])
    with pytest.raises(Exception, match="no device selected"):
        pr.Solver(I, gpus=[])
    S = pr.Solver(I, gpus=[0], nthreads=2)
    bad = pr.Reconstruction(np.zeros((32, 32), np.complex64), np.zeros((32, 32), np.int32))
    with pytest.raises(Exception, match="size differs"):
        S.ER([bad], 1)
