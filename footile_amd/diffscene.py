"""Seeded scenes for the differentiable fill (same SplitMix64 streams as the
rasteriser workloads).

  scene_one_image     one 512x512 image, 256 closed paths of 8 cubics, control
                      points uniform in the canvas, premultiplied RGBA colours
                      with alpha in [0.3, 0.9].                         seed 3
  scene_many_images   64 images of 256x256, each with U[32, 512] paths of
                      U[4, 16] cubics, same colour law.                 seed 4

`perturbed` returns the parameters a target is rendered from: control points
+ N(0, 2 px), colours + N(0, 0.05).
"""
from dataclasses import dataclass

import numpy as np

from .workloads import SplitMix64


@dataclass
class DiffScene:
    n_images: int
    width: int
    height: int
    path_image: np.ndarray   # u32 [n_paths], non-decreasing
    path_cubics: np.ndarray  # u32 [n_paths]
    points: np.ndarray       # f32 [n_points, 2]; per path 1 + 3 * cubics points
    colours: np.ndarray      # f32 [n_paths, 4], premultiplied

    @property
    def n_paths(self):
        return len(self.path_image)


def _colours(rng, n):
    a = 0.3 + 0.6 * rng.uniform(n)
    rgb = [a * rng.uniform(n) for _ in range(3)]
    return np.stack(rgb + [a], axis=1).astype(np.float32)


def scene_one_image(seed=3, width=512, height=512, n_paths=256, cubics=8):
    rng = SplitMix64(seed)
    path_image = np.zeros(n_paths, np.uint32)
    path_cubics = np.full(n_paths, cubics, np.uint32)
    n_pts = int(np.sum(1 + 3 * path_cubics.astype(np.int64)))
    xs = rng.uniform(n_pts) * width
    ys = rng.uniform(n_pts) * height
    points = np.stack([xs, ys], axis=1).astype(np.float32)
    return DiffScene(1, width, height, path_image, path_cubics, points, _colours(rng, n_paths))


def scene_many_images(seed=4, n_images=64, size=256, min_paths=32, max_paths=512, min_cubics=4, max_cubics=16):
    rng = SplitMix64(seed)
    image, cubics = [], []
    for i in range(n_images):
        n = min_paths + rng.below(max_paths - min_paths + 1)
        for _ in range(n):
            image.append(i)
            cubics.append(min_cubics + rng.below(max_cubics - min_cubics + 1))
    path_image = np.asarray(image, np.uint32)
    path_cubics = np.asarray(cubics, np.uint32)
    n_pts = int(np.sum(1 + 3 * path_cubics.astype(np.int64)))
    xs = rng.uniform(n_pts) * size
    ys = rng.uniform(n_pts) * size
    points = np.stack([xs, ys], axis=1).astype(np.float32)
    return DiffScene(n_images, size, size, path_image, path_cubics, points, _colours(rng, len(image)))


def perturbed(scene, seed=7, sigma_px=2.0, sigma_colour=0.05):
    """Parameters the target image is rendered from."""
    rng = SplitMix64(seed)
    pts = scene.points.astype(np.float64) + sigma_px * rng.normal(scene.points.size).reshape(scene.points.shape)
    col = scene.colours.astype(np.float64) + sigma_colour * rng.normal(scene.colours.size).reshape(
        scene.colours.shape)
    return pts.astype(np.float32), col.astype(np.float32)
