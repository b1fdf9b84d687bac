"""CPU checks of the MATLAB binding (no MATLAB here): the MEX gateway type-checks
against include/gpe_hip.h, every command the wrapper classes use exists in the
gateway, and the gateway reaches every solver entry point of the C ABI."""
import glob
import os
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

GATEWAY = os.path.join(ROOT, "gpe-matlab_amd", "mex", "gpe_mex.cpp")
MATLAB = os.path.join(ROOT, "gpe-matlab_amd", "matlab")
HEADER = os.path.join(ROOT, "include", "gpe_hip.h")


def _src():
    return open(GATEWAY).read()


def gateway_commands():
    return set(re.findall(r'cmd == "([a-z_0-9]+)"', _src()))


def declared():
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    return set(re.findall(r"\b(gpe_[a-z_0-9]+)\s*\(", src))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_gateway_type_checks_against_the_c_abi():
    # tests/mex_api/mex.h declares the MEX C API subset the gateway uses; the compiler
    # checks every gpe_* call against the real header (argument count and types)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I", os.path.join(ROOT, "tests", "mex_api"), "-I", os.path.join(ROOT, "include"), GATEWAY],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_gateway_calls_only_declared_entry_points():
    called = set(re.findall(r"\b(gpe_[a-z_0-9]+)\s*\(", _src())) - {"gpe_mex"}
    assert called <= declared(), called - declared()


def test_gateway_reaches_the_solver_entry_points():
    called = set(re.findall(r"\b(gpe_[a-z_0-9]+)\s*\(", _src()))
    for name in ["gpe_create", "gpe_destroy", "gpe_set_params", "gpe_set_potential", "gpe_set_potential_mu",
                 "gpe_set_state", "gpe_get_state", "gpe_set_coupling", "gpe_set_omega", "gpe_split_begin",
                 "gpe_split_outer", "gpe_split_block", "gpe_measure", "gpe_itp", "gpe_itp2", "gpe_sgpe_begin",
                 "gpe_sgpe_outer", "gpe_sgpe_block", "gpe_vtd_set", "gpe_vtd_coeffs", "gpe_detect_core",
                 "gpe_control_set", "gpe_control_forward", "gpe_adjoint_grad", "gpe_rabi_forward", "gpe_rabi_grad",
                 "gpe_batch_create", "gpe_batch_destroy", "gpe_batch_set", "gpe_batch_get_state",
                 "gpe_batch_split_begin", "gpe_batch_split_outer", "gpe_batch_tune", "gpe_tune",
                 "gpe_wisdom_forget", "gpe_wisdom_export", "gpe_wisdom_import", "gpe_last_error"]:
        assert name in called, name


def test_wrapper_classes_use_existing_commands():
    files = glob.glob(os.path.join(MATLAB, "**", "*.m"), recursive=True)
    assert any(f.endswith(os.path.join("@GPEtaskHIP", "solve_split.m")) for f in files)
    assert any(f.endswith(os.path.join("@GPEtaskHIP", "groundstate_itp.m")) for f in files)
    cmds = gateway_commands()
    used = set()
    for f in files:
        used |= set(re.findall(r"gpe_mex\('([a-z_0-9]+)'", open(f).read()))
    assert used and used <= cmds, used - cmds


def test_wrapper_method_signatures_match_the_reference_methods():
    # phi = task.solve_split(ddt,niter_inner,niter_outer)       (lib/@GPEtask/solve_split.m:1)
    # [phi, varargout] = groundstate_itp(task,dt,eps,phi0)       (lib/@GPEtask/groundstate_itp.m:1)
    first = lambda p: open(os.path.join(MATLAB, p)).readline().strip()
    assert first("@GPEtaskHIP/solve_split.m") == "function phi = solve_split(task,ddt,niter_inner,niter_outer)"
    assert first("@GPEtaskHIP/groundstate_itp.m") == "function [phi, varargout] = groundstate_itp(task,dt,eps,phi0)"
    assert first("@GPEtaskSpinorHIP/solve_split.m") == "function phi = solve_split(task,ddt,niter_inner,niter_outer)"
