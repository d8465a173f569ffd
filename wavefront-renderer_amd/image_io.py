"""Output compatibility with the reference renderer: its image files and its stdout report.

* `write_exr` -- what `Image::write` asks OpenImageIO for (src/core/Image.cpp:14-43): an OpenEXR
  file of FLOAT R, G, B channels with ZIP compression and a `Software` attribute. Written here
  directly: single-part scanline layout, 16 scanlines per ZIP chunk, channels stored in name
  order (B, G, R), each chunk byte-split, delta-predicted and deflated as the OpenEXR format
  specifies. `read_exr` reads such files back (NONE, ZIPS and ZIP chunks of HALF or FLOAT
  channels).
* `write_preview_png` -- `Image::writePreview` (Image.cpp:45-71): each channel clamped to
  [0, 1], raised to 1/2.2, scaled by 255, +0.5, truncated to 8 bits; an 8-bit RGB PNG.
* `stats_block`, `render_report` -- `printStatsBlock` (include/core/PrintUtils.h) and the blocks
  `renderScene` prints after a render (src/render/WavefrontRenderer.cpp:321-370), plus main's
  `Rendered:` and `Shading time:` lines (src/main.cpp:203-206), with the labels the reference's
  results tooling matches (scripts/results/parse_results.py:33-45).
"""
from __future__ import annotations

import struct
import zlib
from pathlib import Path

import numpy as np

# ---- OpenEXR -----------------------------------------------------------------------------------
_EXR_MAGIC = 20000630
_EXR_VERSION = 2                      # single-part scanline, short names
_HALF, _FLOAT = 1, 2                  # channel pixel types
_NONE, _ZIPS, _ZIP = 0, 2, 3          # compression
_LINES = {_NONE: 1, _ZIPS: 1, _ZIP: 16}
SOFTWARE = "Wavefront Renderer - MSc Thesis"   # the attribute Image::write sets


def _attr(name: str, kind: str, value: bytes) -> bytes:
    return name.encode() + b"\0" + kind.encode() + b"\0" + struct.pack("<i", len(value)) + value


def _zip_pack(raw: bytes) -> bytes:
    """OpenEXR ZIP/ZIPS chunk: bytes split into even and odd halves, delta-predicted, deflated."""
    a = np.frombuffer(raw, np.uint8)
    t = np.concatenate([a[0::2], a[1::2]])
    d = t.copy()
    d[1:] = ((t[1:].astype(np.int32) - t[:-1].astype(np.int32) + 384) & 0xFF).astype(np.uint8)
    return zlib.compress(d.tobytes())


def _zip_unpack(packed: bytes, size: int) -> bytes:
    d = np.frombuffer(zlib.decompress(packed), np.uint8).astype(np.int64)
    if d.size != size:
        raise ValueError(f"EXR chunk inflates to {d.size} bytes, expected {size}")
    d[1:] -= 128
    t = (np.cumsum(d) & 0xFF).astype(np.uint8)
    out = np.empty(size, np.uint8)
    half = (size + 1) // 2
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return out.tobytes()


def write_exr(path, rgb: np.ndarray, software: str = SOFTWARE) -> None:
    """rgb: (H, W, 3) float image, row 0 at the top."""
    img = np.ascontiguousarray(rgb, np.float32)
    if img.ndim != 3 or img.shape[2] != 3:
        raise ValueError("write_exr expects an (H, W, 3) image")
    h, w = img.shape[:2]
    chlist = b"".join(c.encode() + b"\0" + struct.pack("<iB3xii", _FLOAT, 0, 1, 1)
                      for c in "BGR") + b"\0"
    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    header = (struct.pack("<ii", _EXR_MAGIC, _EXR_VERSION)
              + _attr("channels", "chlist", chlist)
              + _attr("compression", "compression", bytes([_ZIP]))
              + _attr("dataWindow", "box2i", box)
              + _attr("displayWindow", "box2i", box)
              + _attr("lineOrder", "lineOrder", bytes([0]))
              + _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
              + _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
              + _attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
              + _attr("Software", "string", software.encode())
              + b"\0")
    lines = _LINES[_ZIP]
    chunks = []
    for y0 in range(0, h, lines):
        block = img[y0:y0 + lines]                       # (n, W, 3)
        # per scanline: all B values, then G, then R
        raw = np.ascontiguousarray(block[:, :, ::-1].transpose(0, 2, 1)).astype("<f4").tobytes()
        packed = _zip_pack(raw)
        data = packed if len(packed) < len(raw) else raw
        chunks.append(struct.pack("<ii", y0, len(data)) + data)
    offsets, pos = [], len(header) + 8 * len(chunks)
    for c in chunks:
        offsets.append(pos)
        pos += len(c)
    with open(path, "wb") as f:
        f.write(header)
        f.write(struct.pack(f"<{len(offsets)}Q", *offsets))
        for c in chunks:
            f.write(c)


def _parse_header(buf: bytes):
    magic, version = struct.unpack_from("<ii", buf, 0)
    if magic != _EXR_MAGIC:
        raise ValueError("not an OpenEXR file")
    if version & 0xFFFFFF00 & ~0x400:   # tiled / multi-part / deep files are not read here
        raise ValueError(f"unsupported OpenEXR layout (version field {version:#x})")
    pos, attrs = 8, {}
    while buf[pos] != 0:
        end = buf.index(b"\0", pos)
        name = buf[pos:end].decode()
        tend = buf.index(b"\0", end + 1)
        kind = buf[end + 1:tend].decode()
        (size,) = struct.unpack_from("<i", buf, tend + 1)
        attrs[name] = (kind, buf[tend + 5:tend + 5 + size])
        pos = tend + 5 + size
    return attrs, pos + 1


def read_exr_header(path) -> dict:
    """Attribute name -> (type name, raw value bytes)."""
    return _parse_header(Path(path).read_bytes())[0]


def read_exr(path) -> np.ndarray:
    """(H, W, C) float32 image of the file's channels in name order (an RGB file gives R, G, B
    in that order when its channels are B, G, R)."""
    buf = Path(path).read_bytes()
    attrs, pos = _parse_header(buf)
    chl = attrs["channels"][1]
    chans, p = [], 0
    while chl[p] != 0:
        end = chl.index(b"\0", p)
        name = chl[p:end].decode()
        ptype, _, xs, ys = struct.unpack_from("<iB3xii", chl, end + 1)
        if (xs, ys) != (1, 1) or ptype not in (_HALF, _FLOAT):
            raise ValueError(f"unsupported channel {name}")
        chans.append((name, ptype))
        p = end + 1 + 16
    comp = attrs["compression"][1][0]
    if comp not in _LINES:
        raise ValueError(f"unsupported compression {comp}")
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    lines = _LINES[comp]
    n_chunks = (h + lines - 1) // lines
    offsets = struct.unpack_from(f"<{n_chunks}Q", buf, pos)
    widths = [4 if t == _FLOAT else 2 for _, t in chans]
    out = np.zeros((h, w, len(chans)), np.float32)
    for off in offsets:
        yc, size = struct.unpack_from("<ii", buf, off)
        data = buf[off + 8:off + 8 + size]
        n = min(lines, y1 + 1 - yc)
        raw_size = n * w * sum(widths)
        raw = data if size == raw_size or comp == _NONE else _zip_unpack(data, raw_size)
        q = 0
        for r in range(n):
            for c, (name, t) in enumerate(chans):
                dt = "<f4" if t == _FLOAT else "<f2"
                out[yc - y0 + r, :, c] = np.frombuffer(raw, dt, w, q).astype(np.float32)
                q += w * widths[c]
    order = sorted(range(len(chans)), key=lambda c: chans[c][0])
    names = [chans[c][0] for c in order]
    if names == ["B", "G", "R"]:
        return out[:, :, ::-1].copy()
    return out


# ---- PNG preview ---------------------------------------------------------------------------------
def preview_bytes(rgb: np.ndarray) -> np.ndarray:
    """Image::writePreview's gamma-2.2 8-bit conversion, in float32 as the reference computes."""
    v = np.clip(np.asarray(rgb, np.float32), np.float32(0.0), np.float32(1.0))
    g = np.power(v, np.float32(1.0) / np.float32(2.2)) * np.float32(255.0) + np.float32(0.5)
    return g.astype(np.uint8)


def _png_chunk(kind: bytes, data: bytes) -> bytes:
    return (struct.pack(">I", len(data)) + kind + data
            + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF))


def write_png(path, rgb8: np.ndarray) -> None:
    """(H, W, 3) uint8 image as an 8-bit RGB PNG (no filtering)."""
    img = np.ascontiguousarray(rgb8, np.uint8)
    h, w = img.shape[:2]
    rows = b"".join(b"\0" + img[y].tobytes() for y in range(h))
    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n")
        f.write(_png_chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)))
        f.write(_png_chunk(b"IDAT", zlib.compress(rows)))
        f.write(_png_chunk(b"IEND", b""))


def write_preview_png(path, rgb: np.ndarray) -> None:
    write_png(path, preview_bytes(rgb))


# ---- stdout report -------------------------------------------------------------------------------
_RULE = "========================================"


def _num(x: float) -> str:
    """std::to_string(double): fixed, six decimals."""
    return "%f" % x


def stats_block(title: str, lines) -> str:
    """printStatsBlock (include/core/PrintUtils.h)."""
    out = f"\n{_RULE}\n  {title}\n{_RULE}\n"
    out += "".join(f"  {line}\n" for line in lines)
    return out + f"{_RULE}\n\n"


def render_report(stats, width: int, height: int, spp: int, max_depth: int, ray_sort: bool = True,
                  adaptive: bool = False, coherence_stats: int = 0,
                  material_costs=None, texture_costs=None) -> str:
    """The text renderScene and main print after a render (WavefrontRenderer.cpp:321-370,
    main.cpp:203-206) from wf_render_stats. coherence_stats: what the render was asked for
    (1: the Coherence Metrics block, 2: also Memory Coherence Stats). material_costs /
    texture_costs: optional (names, CostTracker) pairs for the cost-tracker blocks
    (cpp:43-65, 372-386)."""
    out = ""
    pipeline = stats.sort_ms + stats.intersect_ms + stats.shade_ms
    overhead = stats.sort_ms / pipeline * 100.0 if pipeline > 0.0 else 0.0
    if ray_sort:
        out += stats_block("Ray Sort Stats", [
            "Total ray sort time (ms)      : " + _num(stats.sort_ms),
            "Total intersect time (ms)     : " + _num(stats.intersect_ms),
            "Total shade time (ms)         : " + _num(stats.shade_ms),
            "Sort overhead (% of pipeline) : " + _num(overhead),
        ])
    weight = stats.shaded_hits
    if weight > 0 and coherence_stats >= 1:
        out += stats_block("Coherence Metrics", [
            "Average material run length  : " + _num(stats.material_run_length),
            "Average texture run length   : " + _num(stats.texture_run_length),
            "Total shaded hits            : " + str(int(weight)),
        ])
        if coherence_stats >= 2:
            out += stats_block("Memory Coherence Stats", [
                "Material ID cache line homogeneity : " + _num(stats.material_homogeneity),
                "Texture  ID cache line homogeneity : " + _num(stats.texture_homogeneity),
            ])
    if adaptive:
        out += stats_block("Adaptive Sampling Stats", [
            "Samples budgeted           : " + str(spp),
            "Samples actually run       : " + str(stats.samples_run),
            "Pixels still active at end : " + str(stats.active_pixels_end),
        ])
    for title, pair in (("Material Cost Tracker Stats", material_costs),
                        ("Texture Cost Tracker Stats", texture_costs)):
        if not pair:
            continue
        names, tracker = pair
        avg, cnt, rel = tracker.query()
        lines = []
        for i, name in enumerate(names):
            if cnt[i] == 0:
                lines.append(f"{name} : no samples yet")
            else:
                lines.append(f"{name} | avg cost: {_num(avg[i])}ns | relative: {_num(rel[i])}"
                             f" | samples: {int(cnt[i])}")
        out += stats_block(title, lines)
    out += f"Rendered: {width}x{height} | {spp} samples | depth {max_depth}\n"
    out += f"Shading time: {stats.shade_ms:g}ms\n"
    return out
