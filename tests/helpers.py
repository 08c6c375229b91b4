"""Shared test helpers: synthetic frames and oracle / device parameter sets."""
import numpy as np

import oracle as O
from nicp_amd import capi, synth

RES = {"160x120": (120, 160), "320x240": (240, 320), "640x480": (480, 640), "1280x960": (960, 1280)}


def cams(rows, cols, min_d=0.5, max_d=4.5):
    fx, fy, cx, cy = synth.camera_for(rows, cols)
    return O.camera(fx, fy, cx, cy, min_d, max_d), capi.camera(fx, fy, cx, cy, min_d, max_d)


def params(**kw):
    return O.stats_params(**kw), capi.stats_params(**kw)


def aparams(**kw):
    return O.align_params(**kw), capi.align_params(**kw)


def identity16():
    return np.eye(4, dtype=np.float32).flatten(order="F")


def angle_dist(T16, U16):
    return synth.pose_error(T16, U16)
