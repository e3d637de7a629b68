"""Camera matrices for the frustum entry points (vx_frustum_from_matrix).

The reference builds the frustum's only input in Camera::update_global_matrix
(src/camera.rs:41-51):

    global_matrix = OPENGL_TO_WGPU_MATRIX * Matrix4::new_perspective(aspect,
                    fov.to_degrees(), near, far) * Matrix4::look_at_rh(eye,
                    eye + target, up)

The C ABI takes that matrix as 16 floats, column-major (nalgebra's storage,
which matrix_to_planes reads, src/frustum_culling.rs:47-85).  These helpers
build it on the host in float64 and round it to float32 once:

  perspective / look_at   nalgebra's new_perspective / look_at_rh
  global_matrix           the reference's product, optionally with its two
                          quirks (`reference_quirks=True`): the field of view
                          passed through to_degrees() into the radians
                          parameter (src/camera.rs:48) and
                          OPENGL_TO_WGPU_MATRIX as written with the row-major
                          Matrix4::new, i.e. the transpose of the usual
                          OpenGL-to-WebGPU depth remap (src/camera.rs:17-23)
  colmajor                4x4 -> the 16 column-major floats the ABI takes
"""
from __future__ import annotations

import numpy as np

# Matrix4::new takes its 16 arguments row by row (src/camera.rs:17-23).
OPENGL_TO_WGPU = np.array([[1.0, 0.0, 0.0, 0.0],
                           [0.0, 1.0, 0.0, 0.0],
                           [0.0, 0.0, 0.5, 0.0],
                           [0.0, 0.0, 0.5, 1.0]])


def perspective(fovy: float, aspect: float, near: float, far: float) -> np.ndarray:
    """Right-handed OpenGL perspective (nalgebra Perspective3 /
    Matrix4::new_perspective(aspect, fovy, near, far)); fovy in radians."""
    f = 1.0 / np.tan(fovy / 2)
    m = np.zeros((4, 4))
    m[0, 0] = f / aspect
    m[1, 1] = f
    m[2, 2] = (far + near) / (near - far)
    m[2, 3] = 2 * far * near / (near - far)
    m[3, 2] = -1
    return m


def look_at(eye, target, up=(0, 1, 0)) -> np.ndarray:
    """Right-handed view matrix (Matrix4::look_at_rh): the camera at `eye`
    looks along -z toward `target`."""
    eye, target, up = (np.asarray(a, np.float64) for a in (eye, target, up))
    f = target - eye
    f = f / np.linalg.norm(f)
    s = np.cross(f, up)
    s = s / np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[:3, 3] = -m[:3, :3] @ eye
    return m


def global_matrix(eye, direction, aspect: float, fov: float, near: float, far: float,
                  up=(0, 1, 0), reference_quirks: bool = False) -> np.ndarray:
    """Camera::update_global_matrix: projection * view toward eye + direction
    (the reference keeps `target` as an offset from the eye).  With
    reference_quirks the reference's own product is reproduced; otherwise fov
    is a plain vertical field of view in radians and no depth remap is
    applied."""
    eye = np.asarray(eye, np.float64)
    if reference_quirks:
        proj = perspective(np.degrees(fov), aspect, near, far)
        return OPENGL_TO_WGPU @ proj @ look_at(eye, eye + np.asarray(direction, np.float64), up)
    return perspective(fov, aspect, near, far) @ look_at(eye, eye + np.asarray(direction, np.float64), up)


def colmajor(m) -> np.ndarray:
    """A 4x4 matrix as the 16 float32 values in column-major order
    (m16[c * 4 + r] = m[r][c])."""
    return np.asarray(m, np.float32).T.reshape(16)
