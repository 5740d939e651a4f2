"""Dataset readers (host code in libse8.so): RAW and TUM formats, ground truth, statuses (CPU).

- The reference's 3-frame ICL-NUIM sequence in TUM layout (test/system/multires_ofusion/dataset:
  depth.txt, rgb.txt, groundtruth.txt and the 16-bit depth PNGs, kept in tests/golden/tum_icl3):
  decoded depth equal to PIL's, metres = stored * float32(1/5000), ground-truth poses equal to the
  file's quaternions turned into rotation matrices (scipy), status error after the last frame.
- PNGs written here with each scanline filter (None, Sub, Up, Average, Paeth) and 8/16 bits decode
  exactly; a corrupted chunk is refused.
- TUM association (a depth frame without an RGB frame within 0.02 s is dropped) and ground-truth
  interpolation (linear position, slerp orientation, compared with scipy's Slerp; a gap above 0.2 s
  drops the frame).
- RAW files written here: depth = stored / 1000 (float32), per-frame ground-truth lines (comment
  lines skipped; a non-unit or non-finite quaternion gives status skip; past the last line, eof),
  restart, and the error status past the last frame.
"""
import struct
import zlib

import numpy as np
import pytest

from supereight2_amd import capi, se

FIX = __file__.rsplit("/", 1)[0] + "/golden/tum_icl3"


def _rot(q_xyzw):
    from scipy.spatial.transform import Rotation

    return Rotation.from_quat(q_xyzw).as_matrix()


def test_icl_sequence_depth_and_ground_truth():
    from PIL import Image

    r = se.Reader("tum", FIX, FIX + "/groundtruth.txt")
    assert (r.width, r.height, r.num_frames, r.has_ground_truth) == (640, 480, 3, True)
    gt = [l.split() for l in open(FIX + "/groundtruth.txt") if not l.startswith("#")]
    inv = np.float32(1.0) / np.float32(5000.0)
    for k in range(3):
        st, raw, scale, divide, T = r.nextDataRaw()
        assert st == se.ReaderStatus.ok and divide == 0 and np.float32(scale) == inv
        png = np.array(Image.open(f"{FIX}/depth/{k + 1}.png")).astype(np.uint16)
        np.testing.assert_array_equal(raw, png)
        v = [float(x) for x in gt[k][1:]]
        np.testing.assert_array_equal(T[:3, 3], np.array(gt[k][1:4], np.float32))
        np.testing.assert_allclose(T[:3, :3], _rot(v[3:]), atol=1e-6)
        np.testing.assert_array_equal(T[3], [0, 0, 0, 1])
    assert r.nextDataRaw()[0] == se.ReaderStatus.error
    r.restart()
    st, depth, T = r.nextData()
    assert st == se.ReaderStatus.ok
    png = np.array(Image.open(f"{FIX}/depth/1.png")).astype(np.uint16)
    np.testing.assert_array_equal(depth, png.astype(np.float32) * inv)
    r.close()


def _png(img, filters):
    """A greyscale PNG of img (uint8 or uint16), scanline k filtered with filters[k % len]."""
    h, w = img.shape
    bpp = img.dtype.itemsize
    rows = img.astype(">u2").tobytes() if bpp == 2 else img.tobytes()
    stride = w * bpp
    raw, prev = bytearray(), bytes(stride)
    for y in range(h):
        line = rows[y * stride:(y + 1) * stride]
        f = filters[y % len(filters)]
        out = bytearray(stride)
        for x in range(stride):
            a = line[x - bpp] if x >= bpp else 0
            b = prev[x]
            c = prev[x - bpp] if x >= bpp else 0
            if f == 0:
                p = 0
            elif f == 1:
                p = a
            elif f == 2:
                p = b
            elif f == 3:
                p = (a + b) // 2
            else:
                pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                p = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
            out[x] = (line[x] - p) & 0xFF
        raw += bytes([f]) + out
        prev = line

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    ihdr = struct.pack(">IIBBBBB", w, h, 8 * bpp, 0, 0, 0, 0)
    return b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr) + chunk(b"IDAT", zlib.compress(bytes(raw))) + chunk(b"IEND", b"")


def _tum_dir(tmp_path, depth_imgs, depth_ts, rgb_ts, gt_lines=None):
    (tmp_path / "depth").mkdir()
    names = []
    for i, img in enumerate(depth_imgs):
        names.append(f"depth/{i}.png")
        (tmp_path / names[-1]).write_bytes(img)
    (tmp_path / "depth.txt").write_text("# timestamp filename\n" + "".join(
        f"{t} {n}\n" for t, n in zip(depth_ts, names)))
    (tmp_path / "rgb.txt").write_text("".join(f"{t} rgb/{i}.png\n" for i, t in enumerate(rgb_ts)))
    if gt_lines is not None:
        (tmp_path / "groundtruth.txt").write_text("# timestamp tx ty tz qx qy qz qw\n" + "".join(gt_lines))
    return str(tmp_path)


@pytest.mark.parametrize("bits", [8, 16])
def test_png_filters(tmp_path, bits):
    rng = np.random.default_rng(bits)
    dt = np.uint16 if bits == 16 else np.uint8
    imgs = [rng.integers(0, 1 << bits, (7, 13)).astype(dt) for _ in range(5)]
    pngs = [_png(imgs[f], [f, (f + 1) % 5, (f + 2) % 5, (f + 3) % 5, (f + 4) % 5]) for f in range(5)]
    d = _tum_dir(tmp_path, pngs, [1, 2, 3, 4, 5], [1, 2, 3, 4, 5])
    r = se.Reader("tum", d)
    for k in range(5):
        st, raw, *_ = r.nextDataRaw(with_pose=False)
        assert st == se.ReaderStatus.ok
        np.testing.assert_array_equal(raw, imgs[k].astype(np.uint16))
    bad = bytearray(pngs[0])
    bad[40] ^= 0xFF  # inside IDAT: the CRC no longer matches
    (tmp_path / "depth" / "1.png").write_bytes(bytes(bad))
    r2 = se.Reader("tum", d)
    assert r2.nextDataRaw(with_pose=False)[0] == se.ReaderStatus.ok  # frame 0 is intact
    assert r2.nextDataRaw(with_pose=False)[0] == se.ReaderStatus.error


def test_tum_association_and_interpolation(tmp_path):
    from scipy.spatial.transform import Rotation, Slerp

    img = _png(np.full((4, 6), 5000, np.uint16), [0])
    # depth 3.0 has no RGB frame within 0.02 s; depth 5.0 lies in a 0.5 s ground-truth gap
    depth_ts = [1.0, 1.1, 2.0, 3.0, 5.0]
    rgb_ts = [1.01, 1.09, 2.0, 3.05, 5.0]
    q0 = Rotation.from_euler("xyz", [0.1, -0.2, 0.3]).as_quat()
    q1 = Rotation.from_euler("xyz", [0.4, 0.1, -0.5]).as_quat()
    gt = [f"1.0 0 0 0 {q0[0]} {q0[1]} {q0[2]} {q0[3]}\n", f"1.2 1 2 3 {q1[0]} {q1[1]} {q1[2]} {q1[3]}\n",
          f"2.0 1 2 3 {q1[0]} {q1[1]} {q1[2]} {q1[3]}\n", f"4.8 0 0 0 0 0 0 1\n", f"5.3 0 0 0 0 0 0 1\n"]
    d = _tum_dir(tmp_path, [img] * 5, depth_ts, rgb_ts, gt)
    r = se.Reader("tum", d, d + "/groundtruth.txt")
    assert r.num_frames == 3  # 1.0, 1.1, 2.0
    st, depth, T = r.nextData()
    np.testing.assert_allclose(T[:3, :3], Rotation.from_quat(q0).as_matrix(), atol=1e-6)
    st, depth, T = r.nextData()
    assert st == se.ReaderStatus.ok and (depth == np.float32(5000) * (np.float32(1) / np.float32(5000))).all()
    np.testing.assert_allclose(T[:3, 3], [0.5, 1.0, 1.5], atol=1e-6)
    ref = Slerp([0, 1], Rotation.from_quat([q0, q1]))(0.5).as_matrix()
    np.testing.assert_allclose(T[:3, :3], ref, atol=2e-6)
    st, depth, T = r.nextData()
    np.testing.assert_allclose(T[:3, 3], [1, 2, 3], atol=1e-6)
    assert r.nextData()[0] == se.ReaderStatus.error


def _raw_file(path, frames, cw=3, ch=2):
    with open(path, "wb") as f:
        for d in frames:
            h, w = d.shape
            f.write(struct.pack("<II", w, h) + d.astype("<u2").tobytes())
            f.write(struct.pack("<II", cw, ch) + bytes(3 * cw * ch))


def test_raw_reader(tmp_path):
    rng = np.random.default_rng(3)
    frames = [rng.integers(0, 65536, (5, 7)).astype(np.uint16) for _ in range(4)]
    _raw_file(tmp_path / "seq.raw", frames)
    q = [0.1, -0.2, 0.3, 0.9]
    qn = np.array(q) / np.linalg.norm(q)
    gt = ["# tx ty tz qx qy qz qw\n", f"0 1 2 3 {qn[0]} {qn[1]} {qn[2]} {qn[3]}\n", "0 0 0 0 0 0 0.5\n",
          "0 0 nan 0 0 0 1\n", "4 5 6 0 0 0 1\n"]
    (tmp_path / "gt.txt").write_text("".join(gt))
    r = se.Reader("raw", tmp_path / "seq.raw", tmp_path / "gt.txt")
    assert (r.width, r.height, r.num_frames) == (7, 5, 4)
    expect = [se.ReaderStatus.ok, se.ReaderStatus.skip, se.ReaderStatus.skip, se.ReaderStatus.ok]
    for k in range(4):
        st, depth, T = r.nextData()
        assert st == expect[k], k
        np.testing.assert_array_equal(depth, frames[k].astype(np.float32) / np.float32(1000.0))
        if k == 0:
            np.testing.assert_allclose(T[:3, 3], [1, 2, 3])
            np.testing.assert_allclose(T[:3, :3], _rot(qn), atol=1e-6)
        if k == 3:
            np.testing.assert_array_equal(T[:3, 3], [4, 5, 6])
    assert r.nextData()[0] == se.ReaderStatus.error
    r.restart()
    st, depth, T = r.nextData()
    assert st == se.ReaderStatus.ok and np.array_equal(T[:3, 3], [1, 2, 3])
    # ground truth shorter than the sequence: eof once its lines run out
    (tmp_path / "gt2.txt").write_text(gt[1])
    r2 = se.Reader("raw", tmp_path / "seq.raw", tmp_path / "gt2.txt")
    assert r2.nextData()[0] == se.ReaderStatus.ok
    assert r2.nextData()[0] == se.ReaderStatus.eof


def test_reader_errors(tmp_path):
    with pytest.raises(capi.Se8Error, match="TUM sequence path"):
        se.Reader("tum", tmp_path / "missing")
    with pytest.raises(capi.Se8Error, match="could not read raw file"):
        se.Reader("raw", tmp_path / "missing.raw")
    (tmp_path / "depth").mkdir()
    (tmp_path / "depth.txt").write_text("1 depth/1.png extra\n")
    (tmp_path / "rgb.txt").write_text("1 rgb/1.png\n")
    with pytest.raises(capi.Se8Error, match="timestamp filename"):
        se.Reader("tum", tmp_path)
