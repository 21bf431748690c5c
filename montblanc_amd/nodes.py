"""Kernel nodes: the v4 node classes, backed by the gfx950 library.

Same class names, lifecycle and solver-array contract as the reference's
nodes (montblanc/impl/rime/v4/gpu/RimeBSqrt.py:151-228,
RimeEBeam.py:373-463, RimeEKBSqrt.py:201-276,
RimeSumCoherencies.py:434-539, v5/gpu/RimeReduction.py:26-64): each reads
the solver's device arrays and writes its outputs back onto the solver.
``RimeFusedForward`` has no reference counterpart: it replaces the four
stage nodes plus the reduction with one fused launch (no jones / B_sqrt
scratch), and is the default pipeline.
"""

import ctypes

import torch

from . import _lib
from .pipeline import Node, Pipeline


def stream_handle(solver, stream):
    """A raw hipStream_t for ``stream`` (None -> torch's current stream)."""
    if stream is None:
        return torch.cuda.current_stream(solver.device).cuda_stream
    if isinstance(stream, torch.cuda.Stream):
        return stream.cuda_stream
    return int(stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() else None


def consts(solver):
    p = solver.exact_property
    return _lib.mbx_consts(p('gauss_scale'), p('two_pi_over_c'), p('sigma_sqrd'))


def beam_extents(solver):
    p = solver.exact_property
    return _lib.mbx_beam_extents(p('beam_ll'), p('beam_lm'), p('beam_lfreq'),
                                 p('beam_ul'), p('beam_um'), p('beam_ufreq'))


def options(solver):
    o = 0
    if solver.use_weight_vector():
        o |= _lib.APPLY_WEIGHTS
    if solver.outputs_residuals():
        o |= _lib.RESIDUAL_OUTPUT
    return o


def suffix(solver):
    return 'f32' if solver.is_float() else 'f64'


class _LibNode(Node):
    def initialise(self, solver, stream=None):
        _lib.load()

    def shutdown(self, solver, stream=None):
        pass

    def pre_execution(self, solver, stream=None):
        pass

    def post_execution(self, solver, stream=None):
        pass


class RimeBSqrt(_LibNode):
    """B_sqrt = sqrt(brightness) (RimeBSqrt.py:151-228)."""

    def execute(self, solver, stream=None):
        s = solver
        d = s.const_data()
        _lib.call('mbx_b_sqrt_' + suffix(s), ctypes.byref(d), ptr(s.stokes), ptr(s.alpha),
                  ptr(s.frequency), ptr(s.ref_frequency), ptr(s.B_sqrt),
                  ctypes.c_void_p(stream_handle(s, stream)))


class RimeEBeam(_LibNode):
    """jones = E beam (RimeEBeam.py:373-463)."""

    def execute(self, solver, stream=None):
        s = solver
        d = s.const_data()
        b = beam_extents(s)
        _lib.call('mbx_e_beam_' + suffix(s), ctypes.byref(d), ptr(s.lm),
                  ptr(s.parallactic_angles), ptr(s.point_errors), ptr(s.antenna_scaling),
                  ptr(s.frequency), ptr(s.E_beam), ptr(s.jones), ctypes.byref(b),
                  ctypes.c_void_p(stream_handle(s, stream)))


class RimeEKBSqrt(_LibNode):
    """jones <- jones . K . B_sqrt (RimeEKBSqrt.py:201-276)."""

    def execute(self, solver, stream=None):
        s = solver
        d = s.const_data()
        k = consts(s)
        _lib.call('mbx_ekb_sqrt_' + suffix(s), ctypes.byref(d), ptr(s.uvw), ptr(s.lm),
                  ptr(s.frequency), ptr(s.B_sqrt), ptr(s.jones), ctypes.byref(k),
                  ctypes.c_void_p(stream_handle(s, stream)))


def _reduce_x2(solver, stream, d=None):
    s = solver
    d = s.const_data() if d is None else d
    k = consts(s)
    nb = _lib.load().mbx_chi2_workspace_bytes()
    ws = s.workspace('chi2', nb)
    x2 = s.x2_device_scalar()
    _lib.call('mbx_chi2_reduce_' + suffix(s), ctypes.byref(d), ptr(s.chi_sqrd_result),
              ptr(x2), ctypes.byref(k), options(s), ptr(ws), nb,
              ctypes.c_void_p(stream_handle(s, stream)))


class RimeSumCoherencies(_LibNode):
    """Coherency sum + gains + chi^2 terms, then X2 (RimeSumCoherencies.py:434-539).

    Like the reference node, X2 is reduced after every execution (the
    reference sums ``chi_sqrd_result`` with gpuarray.sum at :531); on a
    partial source batch only the running sum is written.
    """

    def execute(self, solver, stream=None):
        s = solver
        d = s.const_data()
        k = consts(s)
        h = ctypes.c_void_p(stream_handle(s, stream))
        _lib.call('mbx_sum_coherencies_' + suffix(s), ctypes.byref(d), ptr(s.uvw),
                  ptr(s.gauss_shape), ptr(s.sersic_shape), ptr(s.frequency), ptr(s.antenna1),
                  ptr(s.antenna2), ptr(s.jones), ptr(s.flag), ptr(s.weight_vector),
                  ptr(s.observed_vis), ptr(s.G_term), ptr(s.model_vis), ptr(s.chi_sqrd_result),
                  ctypes.byref(k), options(s), h)
        if d.nsrc.upper_extent >= d.nsrc.global_size:
            _reduce_x2(s, stream, d)


class RimeReduction(_LibNode):
    """X2 = sum of chi_sqrd_result over the active extents (v5/gpu/RimeReduction.py:46-61)."""

    def execute(self, solver, stream=None):
        _reduce_x2(solver, stream)


class RimeFusedForward(_LibNode):
    """The whole forward chain in one launch (no reference counterpart).

    Writes model_vis (or residuals), chi_sqrd_result and X2, exactly what
    the four-node pipeline leaves on the solver, without allocating the
    jones / B_sqrt scratch. If the solver has an ``ekb_vis`` tensor
    registered (``solver.keep_ekb_vis = True``) the pre-gain source sum is
    kept there for the adjoint.
    """

    def execute(self, solver, stream=None):
        s = solver
        d = s.const_data()
        inputs = _lib.mbx_inputs(**{n: (s.tensor(n).data_ptr() if s.tensor(n).numel() else None)
                                    for n in _lib.INPUT_NAMES})
        ekb = s.ekb_vis_tensor() if getattr(s, 'keep_ekb_vis', False) else None
        x2 = s.x2_device_scalar()
        outs = _lib.mbx_outputs(ptr(s.model_vis).value if s.model_vis.numel() else None,
                                s.chi_sqrd_result.data_ptr() if s.chi_sqrd_result.numel() else None,
                                x2.data_ptr(), ekb.data_ptr() if ekb is not None else None)
        nb = _lib.load().mbx_forward_workspace_bytes(ctypes.byref(d), int(s.is_double()))
        ws = s.workspace('forward', nb)
        b = beam_extents(s)
        k = consts(s)
        plan = s.pair_plan()
        _lib.call('mbx_forward_' + suffix(s), ctypes.byref(d), ctypes.byref(inputs),
                  ctypes.byref(outs), ctypes.byref(b), ctypes.byref(k),
                  ctypes.byref(plan[0]) if plan is not None else None, options(s), ptr(ws), nb,
                  ctypes.c_void_p(stream_handle(s, stream)))


def default_pipeline():
    """The fused forward: model visibilities, chi^2 terms and X2."""
    return Pipeline([RimeFusedForward()])


def staged_pipeline():
    """The reference's four-node v4 pipeline (v4/RimeSolver.py:36-40)."""
    return Pipeline([RimeBSqrt(), RimeEBeam(), RimeEKBSqrt(), RimeSumCoherencies()])
