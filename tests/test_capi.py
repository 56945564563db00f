"""CPU checks of the drop-in boundary: the library loads and exports every
entry point that include/claire_amd.h declares (no compute, no GPU)."""
import ctypes
import os

import pytest

import claire_amd


def test_header_declares_the_hot_path():
    names = claire_amd.declared_symbols()
    for required in ("clr_create", "clr_destroy", "clr_hessian_matvec", "clr_solve_state", "clr_solve_adjoint",
                     "clr_solve_inc_state", "clr_solve_inc_adjoint", "clr_apply_inv_reg", "clr_grad", "clr_div",
                     "clr_interp_scalar", "clr_interp_vector", "clr_trajectory", "clr_fft_r2c", "clr_fft_c2r",
                     "clr_evaluate_objective", "clr_evaluate_gradient", "clr_reg_functional"):
        assert required in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(claire_amd.LIB_PATH):
        pytest.fail("libclaire_amd.so not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(claire_amd.LIB_PATH)
    missing = [n for n in claire_amd.declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_last_error_is_empty_string_initially():
    lib = claire_amd.lib()
    assert isinstance(lib.clr_last_error(), bytes)


def test_null_arguments_are_rejected_without_a_gpu():
    lib = claire_amd.lib()
    assert lib.clr_create(None, None, 4, 0, None, None) == 1
    assert b"null" in lib.clr_last_error()


def test_batch_arguments_are_checked_without_a_gpu():
    lib = claire_amd.lib()
    opt = claire_amd.GNOptions()
    assert lib.clr_gn_options_default(ctypes.byref(opt)) == 0
    # an empty batch does nothing
    assert lib.clr_solve_gn_batch(None, None, 0, ctypes.byref(opt), None, None, 1) == 0
    assert lib.clr_solve_gn_batch(None, None, 1, None, None, None, 1) == 1
    assert lib.clr_solve_gn_batch(None, None, -1, ctypes.byref(opt), None, None, 1) == 1
    assert lib.clr_solve_gn_batch(None, None, 0, ctypes.byref(opt), None, None, 0) == 1
    assert b"workers" in lib.clr_last_error()
    ctxs = (ctypes.c_void_p * 2)(None, None)
    assert lib.clr_solve_gn_batch(ctxs, ctxs, 2, ctypes.byref(opt), (claire_amd.GNReport * 2)(), None, 2) == 1
    assert b"null" in lib.clr_last_error()


def test_rk_order_and_graph_setters_reject_a_null_context():
    lib = claire_amd.lib()
    assert lib.clr_set_rk_order(None, 4) == 1
    assert lib.clr_set_matvec_graphs(None, 1) == 1
    assert lib.clr_set_pde_type(None, 0) == 1
    assert lib.clr_set_pde_solver(None, 0) == 1
