"""Helpers shared by the GPU parity tests (HIP library vs CPU oracle)."""
import numpy as np

MISS64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def channels(px):
    px = px.astype(np.int64)
    return np.stack([px & 255, (px >> 8) & 255, (px >> 16) & 255], -1)


def hit_triangle(rec, hit_mode):
    rec = rec.astype(np.uint64)
    if hit_mode == 0:
        return np.where(rec == MISS64, -1, (rec & np.uint64(0xFFFFFFFF)).astype(np.int64))
    return np.where(rec == np.uint64(0xFFFFFFFF), -1, (rec & np.uint64(0xFFFFF)).astype(np.int64))


def hit_distance(rec):
    rec = rec.astype(np.uint64)
    t = (rec >> np.uint64(32)).astype(np.uint32).view(np.float32)
    return np.where(rec == MISS64, np.inf, t)


def compare_frames(eng, ofr, cfg, strict=True):
    """Pass-by-pass comparison.  Returns a report dict; asserts the parity bar:
    identical live-ray and packet counts, hit triangle ids identical on all
    but 1e-4 of rays and only where distances differ by < 1e-5 relative,
    G-buffer and pixels within 1/255 per channel (bit-exact expected)."""
    gp = eng.passes()
    op = ofr.passes
    rep = {"passes": []}
    assert len(gp) == len(op), (len(gp), len(op))
    for i, (g, o) in enumerate(zip(gp, op)):
        assert g["kind"] == o["kind"] and g["slots"] == o["slots"], (i, g, o)
        assert g["rays"] == o["rays"], ("live rays differ", i, g["rays"], o["rays"])
        if cfg.algorithm == 0:
            assert g["chunks"] == o["chunks"], ("packets differ", i, g["chunks"], o["chunks"])
        gh = eng.pass_hits(i)
        oh = o["hits"]
        exact = int((gh != oh).sum())
        entry = dict(kind=g["kind"], rays=g["rays"], mismatched_slots=exact)
        if g["kind"] == 1:
            bad = (gh != oh).sum()
            entry["shadow_flag_mismatch"] = int(bad)
            assert bad <= max(1, 1e-4 * max(1, g["rays"])), (i, int(bad))
        else:
            gt, ot = hit_triangle(gh, cfg.hit_mode), hit_triangle(oh, cfg.hit_mode)
            diff = gt != ot
            entry["hit_id_mismatch"] = int(diff.sum())
            assert diff.sum() <= 1e-4 * max(1, g["rays"]), (i, int(diff.sum()))
            if cfg.hit_mode == 0 and diff.any():
                dg, do = hit_distance(gh[diff]), hit_distance(oh[diff])
                rel = np.abs(dg - do) / np.maximum(np.abs(do), 1e-30)
                assert np.all(rel < 1e-5), rel.max()
        rep["passes"].append(entry)
    gb = eng.gbuffer()
    for k in ("pos", "nrm", "dif", "spec"):
        d = np.abs(gb[k].astype(np.float64) - ofr.gbuffer[k])
        rep["gbuffer_" + k + "_max_abs"] = float(d.max())
        rep["gbuffer_" + k + "_exact"] = bool(np.array_equal(gb[k].view(np.uint32), ofr.gbuffer[k].view(np.uint32)))
    px = eng.pixels()
    dch = np.abs(channels(px) - channels(ofr.pixels))
    rep["pixel_max_channel_diff"] = int(dch.max())
    rep["pixels_exact"] = bool(np.array_equal(px, ofr.pixels))
    assert dch.max() <= 1 or (dch > 1).mean() <= 1e-4, dch.max()
    if strict:
        assert all(p["mismatched_slots"] == 0 for p in rep["passes"]), rep
        assert rep["pixels_exact"], rep
    return rep
