'''
ctypes mirror of `include/tentorch_amd.h` for the tests: the descriptor structs and the entry points
of `libtentorch_hip.so`, so GPU tests can call the C ABI directly with tables they build.
'''

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, 'tentorch_amd', 'libtentorch_hip.so')
HEADER = os.path.join(ROOT, 'include', 'tentorch_amd.h')
MAX_DIMS = 8
TT_F64, TT_F32, TT_C128 = 0, 1, 2


class GssTask(ctypes.Structure):
    _fields_ = [('dst_offset', ctypes.c_int64), ('ndim', ctypes.c_int32), ('first_source', ctypes.c_int32),
                ('num_sources', ctypes.c_int32), ('group', ctypes.c_int32),
                ('extent', ctypes.c_int64 * MAX_DIMS), ('dst_stride', ctypes.c_int64 * MAX_DIMS)]


class GssSource(ctypes.Structure):
    _fields_ = [('src_offset', ctypes.c_int64), ('src_stride', ctypes.c_int64 * MAX_DIMS),
                ('coef_re', ctypes.c_double), ('coef_im', ctypes.c_double)]


class GemmGroup(ctypes.Structure):
    _fields_ = [('c_offset', ctypes.c_int64), ('m', ctypes.c_int64), ('n', ctypes.c_int64),
                ('first_term', ctypes.c_int32), ('num_terms', ctypes.c_int32)]


class GemmTerm(ctypes.Structure):
    _fields_ = [('a_offset', ctypes.c_int64), ('b_offset', ctypes.c_int64), ('k', ctypes.c_int64)]


def load():
    lib = ctypes.CDLL(LIB_PATH)
    vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
    lib.tt_gss_plan_create.argtypes = [ctypes.POINTER(GssTask), i64, ctypes.POINTER(GssSource), i64, ctypes.POINTER(vp)]
    lib.tt_gss_plan_destroy.argtypes = [vp]
    lib.tt_gss_plan_counts.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]
    lib.tt_gss_execute.argtypes = [vp, i32, vp, vp, vp]
    lib.tt_gss_execute_plain.argtypes = [vp, i32, vp, vp, vp]
    lib.tt_gss_execute_typed.argtypes = [vp, i32, i32, vp, vp, vp]
    lib.tt_ggemm_plan_create.argtypes = [ctypes.POINTER(GemmGroup), i64, ctypes.POINTER(GemmTerm), i64, ctypes.POINTER(vp)]
    lib.tt_ggemm_plan_create_config.argtypes = [ctypes.POINTER(GemmGroup), i64, ctypes.POINTER(GemmTerm), i64, i32, ctypes.POINTER(vp)]
    lib.tt_ggemm_config_shape.argtypes = [i32, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32)]
    lib.tt_ggemm_plan_config.argtypes = [vp]
    lib.tt_mlp_param_count.argtypes = [i32, i32]
    lib.tt_mlp_param_count.restype = i64
    lib.tt_mlp_forward.argtypes = [i32, i32, i64, vp, vp, vp, vp]
    lib.tt_mlp_gradient.argtypes = [i32, i32, i64, vp, vp, vp, vp, vp, vp]
    lib.tt_mlp_adam.argtypes = [i64, vp, vp, vp, vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                i32, vp]
    lib.tt_ggemm_plan_destroy.argtypes = [vp]
    lib.tt_ggemm_plan_counts.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i64)]
    lib.tt_ggemm_execute.argtypes = [vp, i32, i32, i32, i32, vp, vp, vp, vp]
    lib.tt_ggemm_execute_plain.argtypes = [vp, i32, vp, vp, vp, vp]
    lib.tt_dot.argtypes = [i32, i64, vp, vp, vp, vp]
    lib.tt_axpby.argtypes = [i32, i64, ctypes.c_double, vp, vp, ctypes.c_double, vp, vp, vp]
    lib.tt_convert.argtypes = [i32, i32, i64, vp, vp, vp]
    lib.tt_multidot.argtypes = [i64, i32, vp, i64, vp, vp, vp]
    lib.tt_multiaxpy.argtypes = [i64, i32, vp, ctypes.c_double, vp, i64, vp, vp]
    lib.tt_lanczos_orthonormalize.argtypes = [i64, i32, vp, i64, vp, vp, vp, vp, vp]
    lib.tt_lanczos_ritz.argtypes = [i32, vp, vp, ctypes.c_double, vp, vp, vp, vp]
    lib.tt_multiaxpy_counted.argtypes = [i64, i32, vp, vp, ctypes.c_double, vp, i64, vp, vp]
    lib.tt_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    lib.tt_launch_count.restype = ctypes.c_int64
    return lib


def error(lib) -> str:
    buf = ctypes.create_string_buffer(1024)
    lib.tt_last_error(buf, 1024)
    return buf.value.decode()


def declared_symbols():
    '''The entry points `include/tentorch_amd.h` declares.'''
    import re
    text = open(HEADER).read()
    return sorted(set(re.findall(r'^\s*(?:int|void)\s+(tt_[a-z0-9_]+)\s*\(', text, re.M)))


def gemm_tables(groups, terms):
    g = (GemmGroup * max(len(groups), 1))()
    for i, (c, m, n, f, k) in enumerate(groups):
        g[i] = GemmGroup(c, m, n, f, k)
    t = (GemmTerm * max(len(terms), 1))()
    for i, (a, b, k) in enumerate(terms):
        t[i] = GemmTerm(a, b, k)
    return g, t


def gss_tables(tasks, sources):
    '''Tables in the engine's inspection format: tasks (dst_offset, ext, dstr, first, count) and
    sources (src_offset, strides, coef).'''
    tt = (GssTask * max(len(tasks), 1))()
    for i, (doff, ext, dstr, first, count) in enumerate(tasks):
        task = GssTask()
        task.dst_offset, task.ndim, task.first_source, task.num_sources, task.group = doff, len(ext), first, count, 0
        for d, (e, s) in enumerate(zip(ext, dstr)):
            task.extent[d], task.dst_stride[d] = e, s
        tt[i] = task
    ss = (GssSource * max(len(sources), 1))()
    for i, (soff, sstr, coef) in enumerate(sources):
        src = GssSource()
        src.src_offset = soff
        for d, s in enumerate(sstr[:MAX_DIMS]):
            src.src_stride[d] = s
        src.coef_re, src.coef_im = coef.real, coef.imag
        ss[i] = src
    return tt, ss
