"""numpy / ctypes views of the C-ABI records (include/diomedes_hip.h).

The structured dtypes are byte-identical to the reference's ``#[repr(C)]``
records (engine/src/render/compute.rs:28-133, engine/src/render/gpu_physics.rs:
18-71), so arrays of them can be handed to the library without conversion.
"""

from __future__ import annotations

import ctypes

import numpy as np

CONTAINER = 0xFFFFFFFF

SPHERE, CUBE, TETRAHEDRON, BOX, HULL = 0, 1, 2, 3, 4
SPRING, ROD, BALL_SOCKET, ROPE = 0, 1, 2, 3
# hard joints: the rod, socket and rope geometry held by a multiplier and a
# ramped penalty per row (the stiffness field is unused)
HARD_ROD, HARD_SOCKET, HARD_ROPE = 4, 5, 6
# bending joints: b's orientation held at a's times a rest rotation (anchor_a
# = the rest quaternion x, y, z, w), softly (stiffness in anchor_b.w) or hard
BEND, HARD_BEND = 7, 8

GPU_BODY = np.dtype(
    [
        ("position", "<f4", (4,)),
        ("orientation", "<f4", (4,)),  # x, y, z, w
        ("tag", "<u4"),
        ("pad", "<u4", (3,)),
        ("corners", "<f4", (4, 4)),
    ],
    align=True,
)
GPU_PAIR = np.dtype([("a", "<u4"), ("b", "<u4")])
GPU_CONTACT_OUT = np.dtype(
    [
        ("valid", "<u4"),
        ("a", "<u4"),
        ("b", "<u4"),
        ("pad", "<u4"),
        ("normal", "<f4", (4,)),
        ("depth", "<f4"),
        ("pad2", "<f4", (3,)),
        ("point_a", "<f4", (4,)),
        ("point_b", "<f4", (4,)),
    ],
    align=True,
)
BODY_STATE = np.dtype(
    [
        ("vel", "<f4", (4,)),
        ("ang", "<f4", (4,)),
        ("prev_vel", "<f4", (4,)),
        ("inv_moment", "<f4", (4,)),
        ("inv_mass", "<f4"),
        ("friction", "<f4"),
        ("pad", "<f4", (2,)),
    ],
    align=True,
)
AVBD_CONTACT = np.dtype(
    [
        ("a", "<u4"),
        ("b", "<u4"),
        ("friction", "<f4"),
        ("pad", "<f4"),
        ("normal", "<f4", (4,)),
        ("tangent1", "<f4", (4,)),
        ("tangent2", "<f4", (4,)),
        ("r_a", "<f4", (4,)),
        ("r_b", "<f4", (4,)),
        ("c0", "<f4", (4,)),
        ("penalty", "<f4", (4,)),
        ("lambda_", "<f4", (4,)),
    ],
    align=True,
)
AVBD_CONSTRAINT = np.dtype(
    [
        ("a", "<u4"),
        ("b", "<u4"),
        ("kind", "<u4"),
        ("pad", "<u4"),
        ("anchor_a", "<f4", (4,)),  # .w = rest / max length
        ("anchor_b", "<f4", (4,)),  # .w = stiffness
    ],
    align=True,
)

assert GPU_BODY.itemsize == 112
assert GPU_PAIR.itemsize == 8
assert GPU_CONTACT_OUT.itemsize == 80
assert BODY_STATE.itemsize == 80
assert AVBD_CONTACT.itemsize == 144
assert AVBD_CONSTRAINT.itemsize == 48


class RunOptions(ctypes.Structure):
    """AvbdRunOptions (engine/src/render/gpu_physics.rs:76-83)."""

    _fields_ = [
        ("dt", ctypes.c_float),
        ("gravity", ctypes.c_float * 3),
        ("alpha", ctypes.c_float),
        ("beta_lin", ctypes.c_float),
        ("penalty_max", ctypes.c_float),
        ("iterations", ctypes.c_uint32),
    ]


class Options(ctypes.Structure):
    """AvbdOptions (engine/src/physics/gpu.rs:127-138)."""

    _fields_ = [
        ("dt", ctypes.c_float),
        ("gravity", ctypes.c_float * 3),
        ("iterations", ctypes.c_uint32),
        ("alpha", ctypes.c_float),
        ("beta_lin", ctypes.c_float),
        ("margin", ctypes.c_float),
        ("penalty_min", ctypes.c_float),
        ("penalty_max", ctypes.c_float),
        ("gamma", ctypes.c_float),
        ("friction", ctypes.c_float),
    ]


class Container(ctypes.Structure):
    _fields_ = [
        ("center", ctypes.c_float * 3),
        ("rotation", ctypes.c_float * 4),
        ("half_extent", ctypes.c_float),
        ("enabled", ctypes.c_uint32),
    ]


class Limits(ctypes.Structure):
    _fields_ = [
        ("max_bodies", ctypes.c_uint32),
        ("max_pairs", ctypes.c_uint32),
        ("max_contacts", ctypes.c_uint32),
        ("max_constraints", ctypes.c_uint32),
        ("max_hull_vertices", ctypes.c_uint32),
    ]


class StepOpts(ctypes.Structure):
    _fields_ = [("contact_lag", ctypes.c_uint32), ("mode", ctypes.c_uint32), ("profile", ctypes.c_uint32)]


class StepStats(ctypes.Structure):
    _fields_ = [
        ("steps", ctypes.c_uint32),
        ("pairs", ctypes.c_uint32),
        ("contacts", ctypes.c_uint32),
        ("csr_nnz", ctypes.c_uint32),
        ("narrow_contacts", ctypes.c_uint32),
        ("dynamic_bodies", ctypes.c_uint32),
        ("sum_pairs", ctypes.c_uint64),
        ("sum_contacts", ctypes.c_uint64),
        ("sum_csr_nnz", ctypes.c_uint64),
        ("ms_total", ctypes.c_float),
        ("ms_avbd", ctypes.c_float),
        ("ms_init", ctypes.c_float),
        ("ms_primal", ctypes.c_float),
        ("ms_dual", ctypes.c_float),
        ("ms_recover", ctypes.c_float),
        ("n_init", ctypes.c_uint32),
        ("n_primal", ctypes.c_uint32),
        ("n_dual", ctypes.c_uint32),
        ("n_recover", ctypes.c_uint32),
        ("colours", ctypes.c_uint32),
        ("colour_rounds", ctypes.c_uint32),
    ]


assert ctypes.sizeof(StepStats) == 96


class Tuning(ctypes.Structure):
    _fields_ = [("primal_lanes", ctypes.c_uint32), ("broad_lanes", ctypes.c_uint32),
                ("resort_interval", ctypes.c_uint32), ("fuse_dual", ctypes.c_uint32),
                ("adjoint_lanes", ctypes.c_uint32), ("overlap_broad", ctypes.c_uint32),
                ("primal_order", ctypes.c_uint32), ("scene_solve", ctypes.c_uint32),
                ("colour_lanes", ctypes.c_uint32), ("primal_packed", ctypes.c_uint32)]


def f32(x) -> float:
    """Round a Python float to the nearest float32 (what the ABI receives)."""
    return float(np.float32(x))
