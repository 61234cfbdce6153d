"""Speaker pressure source on the device (SpeakerPressureSource.cpp:91-113):
a pre-step pass writes the speakers' S dt over the cells within 8 widths of
them, and every tile with a source adds it; parity with the FP64 oracle
(listener relative L2 <= 1e-4) for a speaker alone and beside a rigid
sphere, every tile plan identical, and the XML path with a WAV file."""
import os
import wave

import numpy as np
import pytest

from acoustics_amd import scenes
from acoustics_amd.solver import Solver, rel_l2
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("with_sphere", [False, True])
def test_speaker_matches_oracle(hip_lib, with_sphere):
    sb, steps = scenes.speaker_scene(seed=15, with_sphere=with_sphere)
    print(f"case: speaker sphere={with_sphere} divisions={sb.divisions} steps={steps}", flush=True)
    ref = Oracle(sb, reclassify_each_step=False).step(steps)
    with Solver(sb) as s:
        s.step(steps)
        got = s.series()
    err = rel_l2(got, ref)
    print("speaker rel L2", err, "peak", np.abs(ref).max())
    assert np.abs(ref).max() > 0
    assert err <= 1e-4


def test_speaker_sample_ties_follow_the_accumulated_clock(hip_lib):
    """A speaker on a cell centre with the sample period dt / 2 and samples of
    alternating sign: at every step the sample index sits on an integer up to
    the last bits of the time, and a clock computed as t_start + s dt picks
    the other sample on about half of the steps (184 of 400 here).  The
    device accumulates the time as the reference and the oracle do."""
    sb, steps = scenes.speaker_tie_scene()
    print(f"case: speaker ties divisions={sb.divisions} steps={steps}", flush=True)
    ref = Oracle(sb, reclassify_each_step=False).step(steps)
    with Solver(sb, steps_per_graph=16) as s:
        s.step(steps)
        got = s.series()
    err = rel_l2(got, ref)
    print("speaker ties rel L2", err, "peak", np.abs(ref).max())
    assert np.abs(ref).max() > 0
    assert err <= 1e-4


def test_speaker_tile_plans_agree(hip_lib):
    sb, _ = scenes.speaker_scene(seed=16, n=40, with_sphere=True)
    print(f"case: speaker plans divisions={sb.divisions}", flush=True)
    with Solver(sb, tile_config=0) as a:
        a.step(80)
        ea = a.pressure()
    for tile in range(11):
        with Solver(sb, tile_config=tile, steps_per_graph=16 if tile % 2 else 0) as b:
            b.step(80)
            assert np.array_equal(ea, b.pressure()), tile


def test_xml_scene_with_speaker(hip_lib, tmp_path):
    fs = 44100
    x = (scenes.tone_burst(fs=fs) * 32767).astype("<i2")
    with wave.open(str(tmp_path / "tone.wav"), "wb") as w:
        w.setnchannels(1)
        w.setsampwidth(2)
        w.setframerate(fs)
        w.writeframes(x.tobytes())
    text = open(os.path.join(ROOT, "tests", "golden", "scene_sphere.xml")).read()
    spk = ('<speaker_pressure_source start_time="0.0001" width_space="0.008" speaker_file="tone.wav" '
           'position_x="0.02" position_y="0.0" position_z="0.0"/>')
    text = (text.replace("<pressure_source>", "<pressure_source>" + spk, 1) if "<pressure_source>" in text
            else text.replace("</scene>", "<pressure_source>" + spk + "</pressure_source></scene>"))
    q = tmp_path / "spk.xml"
    q.write_text(text)
    print(f"case: speaker xml {q}", flush=True)
    with Solver.from_xml(str(q)) as s:
        s.step(200)
        got = s.series()
    # against the same scene without the speaker, the tone adds its own
    # (small, about 1e-5 Pa at these listeners) signal
    quiet = text.replace(spk, "")
    q2 = tmp_path / "quiet.xml"
    q2.write_text(quiet)
    with Solver.from_xml(str(q2)) as s:
        s.step(200)
        base = s.series()
    assert got.shape == base.shape == (200, base.shape[1])
    assert np.all(np.isfinite(got))
    assert np.abs(got - base).max() > 1e-6
