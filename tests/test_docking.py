"""run_unidock (unidock_tools.docking): the reference's Python docking entry
(unidock_tools/modules/docking/unidock.py:155-201) run in-process.  Argument
checks run on the CPU; the docking runs are GPU tests on the reference's own
fixtures with its acceptance range (test_unidock.py, test_unidock_sdf.py:
best score in [-20, -1] kcal/mol)."""
import os

import pytest

from unidock_tools import docking

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_search_modes_match_the_cli_presets():
    assert docking.SEARCH_MODES["fast"] == (128, 20)
    assert docking.SEARCH_MODES["balance"] == (384, 40)
    assert docking.SEARCH_MODES["detail"] == (512, 40)


def test_unbuilt_options_are_refused(tmp_path):
    args = dict(receptor=os.path.join(GOLDEN, "1iep_receptor.pdbqt"),
                ligands=[os.path.join(GOLDEN, "1iep_ligand.pdbqt")], output_dir=tmp_path,
                center_x=15.19, center_y=53.903, center_z=16.917)
    with pytest.raises(NotImplementedError):
        docking.run_unidock(scoring="dkoes", **args)
    with pytest.raises(NotImplementedError):
        docking.run_unidock(bias_file="x.bpf", **args)
    with pytest.raises(ValueError):
        docking.run_unidock(search_mode="quick", **args)


def test_global_steps_heuristic():
    class L:
        n_atoms, n_tors = 24, 5
    # 70 * 3 * (50 + 24 + 10 * 11) / 2 = 19320, capped by max_step
    assert docking.global_steps_for([L()], 20) == 20
    assert docking.global_steps_for([L()], 0) == 19320


@pytest.mark.gpu
def test_run_unidock_on_reference_fixtures(tmp_path):
    ligs = [os.path.join(GOLDEN, "1iep_ligand.pdbqt")]
    out, scores = docking.run_unidock(os.path.join(GOLDEN, "1iep_receptor.pdbqt"), ligs, tmp_path,
                                      15.19, 53.903, 16.917, 20, 20, 20, search_mode="fast", num_modes=9,
                                      refine_step=3)
    print(f"[gpu] run_unidock 1iep: {scores}", flush=True)
    assert [p.name for p in out] == ["1iep_ligand_out.pdbqt"]
    assert 1 <= len(scores[0]) <= 9 and -20 <= scores[0][0] <= -1
    assert scores[0] == sorted(scores[0])
    out, scores = docking.run_unidock(os.path.join(GOLDEN, "1a30_protein.pdbqt"),
                                      [os.path.join(GOLDEN, "1a30_ligand.sdf")], tmp_path,
                                      8.729, 25.62, 4.682, 19.12, 16.56, 18.65, scoring="vinardo",
                                      search_mode="fast", num_modes=1)
    print(f"[gpu] run_unidock 1a30 vinardo: {scores}", flush=True)
    assert [p.name for p in out] == ["1a30_ligand_out.sdf"] and len(scores[0]) == 1
    assert -20 <= scores[0][0] <= -1
    assert docking.read_scores(out[0]) == scores[0]
    # score only: the input poses
    inp, sc = docking.run_unidock(os.path.join(GOLDEN, "1iep_receptor.pdbqt"), ligs, tmp_path,
                                  15.19, 53.903, 16.917, 20, 20, 20, score_only=True)
    txt = (tmp_path / "scores.txt").read_text().split()
    assert txt[0] == "REMARK" and txt[1] == "1iep_ligand.pdbqt" and float(txt[2]) == pytest.approx(sc[0][0], abs=1e-3)
    assert -20 <= sc[0][0] <= -1  # the crystal pose scores well


@pytest.mark.gpu
def test_run_unidock_ad4(tmp_path):
    """scoring="ad4": AD4.2 maps built on the GPU from the receptor (the
    reference runs AutoGrid for them), docking, AD4.2 final scores."""
    ligs = [os.path.join(GOLDEN, "1iep_ligand.pdbqt")]
    out, scores = docking.run_unidock(os.path.join(GOLDEN, "1iep_receptor.pdbqt"), ligs, tmp_path,
                                      15.19, 53.903, 16.917, 20, 20, 20, scoring="ad4", search_mode="fast",
                                      num_modes=9)
    print(f"[gpu] run_unidock 1iep ad4: {scores}", flush=True)
    assert [p.name for p in out] == ["1iep_ligand_out.pdbqt"]
    assert 1 <= len(scores[0]) <= 9 and -20 <= scores[0][0] <= -1
    assert scores[0] == sorted(scores[0])
    inp, sc = docking.run_unidock(os.path.join(GOLDEN, "1iep_receptor.pdbqt"), ligs, tmp_path,
                                  15.19, 53.903, 16.917, 20, 20, 20, scoring="ad4", score_only=True)
    print(f"[gpu] run_unidock 1iep ad4 score_only: {sc}", flush=True)
    assert -20 <= sc[0][0] <= -1  # the crystal pose scores well
