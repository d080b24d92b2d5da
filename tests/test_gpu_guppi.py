"""GUPPI RAW reader and writer (reference src/modules/guppi/reader.cc,
writer.cc, python/blade/ext_guppi.cc).

Files are written here by an independent Python restatement of the format
(80-byte header cards closed by END, DIRECTIO padding to 512 bytes, data
[NANTS][OBSNCHAN/NANTS][time][pol][re, im] int8) and read back bit for bit;
the writer's files are parsed the same way.  The step timestamps follow the
reference's call (reader.cc:128-139) with the guppirawc99 date routine
restated (unix = SYNCTIME + packet index * TBIN * samples per block /
PIPERBLK, JD = unix / 86400 + 2440587.5): that routine lives in an
un-vendored library, so the timestamp is pinned to this restatement only.
GPU-marked: the modules' host tensors are pinned (hipHostMalloc).
"""

import numpy as np
import pytest

import blade_amd as bl
import oracle as O

pytestmark = pytest.mark.gpu


def card(key, value):
    if isinstance(value, str):
        value = f"'{value}'"
    return f"{key:<8.8}= {str(value):<70.70}".encode()


def write_guppi(path, blocks, cards, directio):
    """blocks: list of int8 arrays [A, F, T, P, 2]; cards: list of (key, value) per block."""
    with open(path, "wb") as fh:
        for data, extra in zip(blocks, cards):
            A, F, T, P, _ = data.shape
            head = [card("BLOCSIZE", data.nbytes), card("OBSNCHAN", A * F), card("NANTS", A),
                    card("NPOL", 4 if P == 2 else P), card("NBITS", 8), card("DIRECTIO", int(directio))]
            head += [card(k, v) for k, v in extra]
            head.append(f"{'END':<80}".encode())
            h = b"".join(head)
            if directio:
                h += b" " * (-len(h) % 512)
            fh.write(h)
            fh.write(np.ascontiguousarray(data).tobytes())
            if directio:
                fh.write(b"\0" * (-data.nbytes % 512))


def parse_guppi(path):
    out = []
    raw = open(path, "rb").read()
    pos = 0
    while pos < len(raw):
        cards = {}
        while True:
            c = raw[pos:pos + 80].decode()
            pos += 80
            if c.startswith("END"):
                break
            cards[c[:8].strip()] = c[10:].strip().strip("'").strip()
        if int(cards.get("DIRECTIO", 0)):
            pos += -pos % 512
        n = int(cards["BLOCSIZE"])
        out.append((cards, raw[pos:pos + n]))
        pos += n + (-n % 512 if int(cards.get("DIRECTIO", 0)) else 0)
    return out


def make_recording(tmp_path, A=3, F=4, T=64, nblocks=(2, 1), directio=True, seed=0):
    rng = np.random.default_rng(seed)
    blocks, metas = [], []
    k = 0
    for fi, nb in enumerate(nblocks):
        fb, fm = [], []
        for _ in range(nb):
            fb.append(rng.integers(-128, 128, size=(A, F, T, 2, 2), dtype=np.int8))
            fm.append([("CHAN_BW", 0.5), ("OBSFREQ", 6500.125), ("SCHAN", 352), ("SYNCTIME", 1649366400),
                       ("PIPERBLK", 16), ("PKTIDX", 1000 + 16 * k), ("DUT1", -0.125)])
            k += 1
        write_guppi(tmp_path / f"rec.{fi:04d}.raw", fb, fm, directio)
        blocks += fb
        metas += fm
    return np.concatenate(blocks, axis=2), metas


def reader(path, nt, nf, na):
    cfg = {"filepath": str(path), "step_number_of_time_samples": nt,
           "step_number_of_frequency_channels": nf, "step_number_of_aspects": na}
    return bl.module(bl.guppi_reader, cfg, (), ot=bl.ci8)


def expected_jd(meta, block_index, T):
    piperblk = meta["PIPERBLK"]
    pkt = meta["PKTIDX"] + int((block_index + 0.5 * piperblk) * T)
    unix = meta["SYNCTIME"] + pkt * ((1.0 / (meta["CHAN_BW"] * 1e6)) * T / piperblk)
    return unix / 86400.0 + 2440587.5


@pytest.mark.parametrize("directio", [True, False])
def test_reader_whole_blocks_across_files(tmp_path, directio):
    data, metas = make_recording(tmp_path, directio=directio)
    r = reader(tmp_path / "rec.0000.raw", 0, 0, 0)      # 0 = the file's own shape
    shape = r.get_total_output_buffer_shape()
    assert (shape.number_of_aspects, shape.number_of_frequency_channels,
            shape.number_of_time_samples, shape.number_of_polarizations) == (3, 4, 192, 2)
    assert r.get_number_of_steps() == 3
    assert r.get_channel_bandwidth() == 0.5e6
    assert r.get_total_bandwidth() == 4 * 0.5e6
    assert r.get_channel_start_index() == 352
    assert r.get_observation_frequency() == 6500.125e6
    for s in range(3):
        assert r.process(s) == bl.result.success
        assert np.array_equal(r.get_step_output_buffer().as_numpy(), data[:, :, 64 * s:64 * (s + 1)])
        meta = dict(metas[s])
        assert r.get_step_output_julian_date().as_numpy()[0] == pytest.approx(expected_jd(meta, s, 64), abs=1e-9)
        assert r.get_step_output_dut().as_numpy()[0] == -0.125
    assert r.process(3) == bl.result.pipeline_exhausted


def test_reader_partial_steps_walk_time_then_channels_then_aspects(tmp_path):
    data, _ = make_recording(tmp_path, nblocks=(3,), directio=False, seed=1)
    r = reader(tmp_path / "rec.0000.raw", 48, 2, 1)
    # 192 samples hold 4 whole steps of 48 (the last crosses no block end);
    # steps spanning block boundaries (48..96, 96..144) read from two blocks.
    expect = []
    for a in range(3):
        for f0 in (0, 2):
            for t0 in range(0, 192 - 48 + 1, 48):
                expect.append(data[a:a + 1, f0:f0 + 2, t0:t0 + 48])
    assert r.get_number_of_steps() == len(expect)
    for s, e in enumerate(expect):
        assert r.process(s) == bl.result.success, s
        assert np.array_equal(r.get_step_output_buffer().as_numpy(), e), s
    assert r.process(len(expect)) == bl.result.pipeline_exhausted


def test_reader_rejects_bad_files(tmp_path):
    (tmp_path / "junk.raw").write_bytes(b"NOT A GUPPI FILE" * 10)
    with pytest.raises(RuntimeError):
        reader(tmp_path / "junk.raw", 0, 0, 0)
    data, _ = make_recording(tmp_path, nblocks=(1,))
    with pytest.raises(RuntimeError):
        reader(tmp_path / "rec.0000.raw", 64, 3, 1)     # 3 channels do not tile 4


@pytest.mark.parametrize("dtype,nbits", [("cf32", 32), ("cf16", 16)])
@pytest.mark.parametrize("directio", [True, False])
def test_writer_blocks_and_header(tmp_path, dtype, nbits, directio):
    rng = np.random.default_rng(5)
    shape = (2, 3, 16, 2)
    buf = bl.array_tensor(shape, dtype=getattr(bl, dtype), device=bl.cpu)
    cfg = {"filepath": str(tmp_path / "out"), "directio": directio, "input_frequency_batches": 1}
    w = bl.module(bl.guppi_writer, cfg, buf, it=getattr(bl, dtype))
    w.header_put("OBSFREQ", 1420.0)
    w.header_put("SCHAN", 7)
    frames = []
    for k in range(3):
        v = rng.normal(size=shape) + 1j * rng.normal(size=shape)
        if dtype == "cf32":
            v = v.astype(np.complex64)
        else:           # CF16 host tensors are float16 pairs (trailing axis of 2)
            v = np.stack([v.real, v.imag], axis=-1).astype(np.float16)
        np.copyto(buf.as_numpy(), v)
        frames.append(v.copy())
        assert w.process(k) == bl.result.success
    del w
    blocks = parse_guppi(tmp_path / "out.0000.raw")
    assert len(blocks) == 3
    for k, (cards, payload) in enumerate(blocks):
        assert int(cards["BLOCSIZE"]) == frames[k].nbytes
        assert int(cards["OBSNCHAN"]) == 6 and int(cards["NANTS"]) == 2 and int(cards["NBEAM"]) == 2
        assert int(cards["NPOL"]) == 2 and int(cards["NBITS"]) == nbits
        assert int(cards["DIRECTIO"]) == int(directio)
        assert cards["DATATYPE"] == "FLOAT"
        assert float(cards["OBSFREQ"]) == 1420.0 and int(cards["SCHAN"]) == 7
        assert np.array_equal(np.frombuffer(payload, dtype=frames[k].dtype).reshape(frames[k].shape),
                              frames[k])
    # reference writer.cc:59: PKTIDX = samples * (blocks written before the previous one)
    assert "PKTIDX" not in blocks[0][0]
    assert [int(b[0]["PKTIDX"]) for b in blocks[1:]] == [0, 16]


def test_writer_rejects_frequency_batches(tmp_path):
    buf = bl.array_tensor((1, 2, 8, 2), dtype=bl.cf32, device=bl.cpu)
    cfg = {"filepath": str(tmp_path / "b"), "directio": False, "input_frequency_batches": 2}
    with pytest.raises(RuntimeError):
        bl.module(bl.guppi_writer, cfg, buf, it=bl.cf32)


def test_reader_feeds_mode_b_pipeline(tmp_path):
    # Ingest: GUPPI blocks -> pinned host step buffer -> H2D on the runner's
    # stream -> Mode-B beams, checked against the oracle per step.
    A, F, T = 4, 8, 512
    data, _ = make_recording(tmp_path, A=A, F=F, T=T, nblocks=(2,), seed=3)
    data = np.clip(data, -127, 127)       # the Mode-B input range
    rng = np.random.default_rng(9)
    w = (rng.normal(size=(2, A, F, 1, 2)) + 1j * rng.normal(size=(2, A, F, 1, 2))).astype(np.complex64)
    write_guppi(tmp_path / "ing.0000.raw", [data[:, :, :T], data[:, :, T:]],
                [[("CHAN_BW", 0.5)], [("CHAN_BW", 0.5)]], True)
    r = reader(tmp_path / "ing.0000.raw", T, F, A)

    @bl.runner
    class Ingest:
        def __init__(self):
            self.input.buf = bl.array_tensor((A, F, T, 2), dtype=bl.ci8)
            self.input.phasors = bl.phasor_tensor((2, A, F, 1, 2), dtype=bl.cf32)
            self.output.buf = bl.array_tensor((2, F, T, 2), dtype=bl.cf32)
            cfg = {"enable_incoherent_beam": False, "enable_incoherent_beam_sqrt": False}
            self.module.bf = bl.module(bl.beamformer, cfg, (self.input.buf, self.input.phasors),
                                       it=bl.ci8, ot=bl.cf32, telescope=bl.ata)

        def transfer_in(self, buf, phasors):
            self.copy(self.input.buf, buf)
            self.copy(self.input.phasors, phasors)

        def transfer_out(self, buf):
            self.copy(self.output.buf, self.module.bf.get_output())
            self.copy(buf, self.output.buf)

    p = Ingest()
    hph = bl.phasor_tensor((2, A, F, 1, 2), dtype=bl.cf32, device=bl.cpu)
    np.copyto(hph.as_numpy(), w)
    hout = bl.array_tensor((2, F, T, 2), dtype=bl.cf32, device=bl.cpu)
    for s in range(2):
        assert r.process(s) == bl.result.success
        assert p(r.get_step_output_buffer(), hph, hout)
        ref = O.beamformer_ata(data[:, :, s * T:(s + 1) * T], w)
        assert O.norm_rel_err(hout.as_numpy(), ref) <= 1e-5, s
