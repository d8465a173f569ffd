"""Progress protocol of the reference renderer, which its GUI follows: every `progressInterval`
samples the running average is written as the gamma-2.2 preview PNG and `Sample: i/N` is printed
(src/render/WavefrontRenderer.cpp:297-311); the preview sits beside the output as
`<output stem>_preview.png` (src/main.cpp:25-33) and is removed when the render ends
(WavefrontRenderer.cpp:388-393). The GUI reads the `Sample:` lines (gui/render/worker.py:131-144)
and polls the PNG.

Here a render of N samples runs as consecutive wf_render calls of `progress_interval` samples
each, the k-th starting at RNG sample index k * progress_interval (wf_render_params
.first_sample), so together they trace exactly the paths of one N-sample render; the running
image is the sample-weighted mean of the chunks (equal to the one-call image up to the order of
the float additions). Adaptive sampling keeps per-pixel state across samples and is rendered in
one call (no intermediate previews).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Callable, Optional

import numpy as np

from . import render
from .image_io import write_preview_png


def preview_path_for(output_path: str) -> str:
    """derivePreviewPath (src/main.cpp:25-33)."""
    dot = output_path.rfind(".")
    return (output_path[:dot] if dot >= 0 else output_path) + "_preview.png"


def render_progressive(scene, params, preview_path: Optional[str] = None,
                       progress_interval: int = 4,
                       callback: Optional[Callable[[int, int], None]] = None,
                       echo: Optional[Callable[[str], None]] = print):
    """Renders params (spp samples) with previews every progress_interval samples. Returns
    (image HxWx3 float32, per-pixel sample counts, list of per-chunk wf_render_stats)."""
    spp = int(params.spp)
    if params.adaptive or progress_interval <= 0 or spp <= progress_interval:
        img, cnt, st = render(scene, params)
        return img, cnt, [st]
    h, w = params.height, params.width
    total = np.zeros((h, w, 3), np.float64)
    stats = []
    done = 0
    try:
        while done < spp:
            chunk = min(progress_interval, spp - done)
            p = type(params)()
            C.memmove(C.byref(p), C.byref(params), C.sizeof(params))
            p.spp = chunk
            p.first_sample = params.first_sample + done
            img, _, st = render(scene, p)
            total += img.astype(np.float64) * chunk
            stats.append(st)
            done += chunk
            if done % progress_interval == 0 or done == spp:
                running = (total / done).astype(np.float32)
                if preview_path:
                    write_preview_png(preview_path, running)
                    if echo:
                        echo(f"Sample: {done}/{spp}")
                if callback:
                    callback(done, spp)
    finally:
        if preview_path and Path(preview_path).exists():
            try:
                os.remove(preview_path)
            except OSError as e:   # the reference warns and goes on (cpp:390-392)
                if echo:
                    echo(f"WARNING: Could not delete preview: {preview_path} ({e})")
    counts = np.full((h, w), spp, np.int32)
    return (total / spp).astype(np.float32), counts, stats
