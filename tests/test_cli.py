"""Command-line front ends on the CPU: word-shape signatures, tree rendering (debinarization), and
CompileGrammar's grz output."""
import subprocess
import sys
import zipfile

import numpy as np

from puck_amd import cli


def test_unknown_word_signatures():
    assert cli.signature("Xqzt", 0, False) == "UNK-INITC"
    assert cli.signature("Xqzt", 0, True) == "UNK-INITC-KNOWNLC"
    assert cli.signature("Xqzt", 3, False) == "UNK-CAPS"
    assert cli.signature("Xyzzy", 0, False) == "UNK-INITC-y"
    assert cli.signature("flurbing", 2, False) == "UNK-LC-ing"
    assert cli.signature("re-flurbs", 2, False) == "UNK-LC-DASH-s"
    assert cli.signature("1987", 1, False) == "UNK-NUM"
    assert cli.signature("A4-B", 1, False) == "UNK-CAPS-NUM-DASH"


def test_word_ids_back_off_to_known_signatures(wsj1):
    ids = cli.word_ids(wsj1, ["The", "zorblaxing", "of", "Qwertz"])
    assert ids is not None and len(ids) == 4
    assert ids[2] == wsj1.word_id("of")
    assert ids[1] == wsj1.word_id("UNK-LC-ing")


def test_render_splices_binarization_nodes():
    # (ROOT (S (NP (DT the) (NN dog)) (VP (VBZ barks)))) with an @S intermediate:
    # nodes are (top, bot, width) in pre-order; names index the arrays below.
    nt = ["ROOT", "S", "@S", "NP", "VP"]
    tt = ["DT", "NN", "VBZ"]
    nodes = np.array([
        [0, 2, 3, 0],    # ROOT over @S (3 words)
        [-1, 3, 2, 0],   # NP (2 words)
        [-1, 0, 1, 0],   # DT the
        [-1, 1, 1, 0],   # NN dog
        [4, 2, 1, 0],    # VP over VBZ barks
    ])
    out = cli.render(nodes, ["the", "dog", "barks"], nt, tt)
    assert out == "(ROOT (NP (DT the) (NN dog)) (VP (VBZ barks)))"


def test_compile_writes_grz(tmp_path, puck):
    out = tmp_path / "g.grz"
    prefix = puck.grammar_path("wsj_0") + ":" + puck.grammar_path("wsj_1")
    subprocess.run([sys.executable, "-m", "puck_amd.cli", "compile", "--textGrammarPrefix", prefix,
                    "--grammar", str(out), "--viterbi", "false"], check=True, cwd=str(tmp_path.parent),
                   env={"PYTHONPATH": cli.os.path.dirname(cli.os.path.dirname(cli.__file__)), "PATH": "/usr/bin"})
    with zipfile.ZipFile(out) as z:
        assert z.read("parser-0/semiring") == b"Viterbi"
        assert z.read("parser-1/semiring") == b"Real"
    cli.main(["compile", "--textGrammarPrefix", prefix, "--grammar", str(out), "--viterbi", "false",
              "--logsum", "true"])
    with zipfile.ZipFile(out) as z:
        assert z.read("parser-1/semiring") == b"LogSum"


def test_compile_reproject(tmp_path, puck):
    out = tmp_path / "r.grz"
    prefix = puck.grammar_path("wsj_0") + ":" + puck.grammar_path("wsj_1")
    cli.main(["compile", "--textGrammarPrefix", prefix, "--grammar", str(out), "--reproject", "true"])
    c, f = puck.Grammar.load_grz(str(out))
    rc, rf = puck.Grammar.reproject(puck.Grammar.load_text(puck.grammar_path("wsj_0")),
                                    puck.Grammar.load_text(puck.grammar_path("wsj_1")))
    assert f.coarse_names() == rf.coarse_names() and f.coarse_names()[1] == "S_0"
    assert np.array_equal(f.rules(0)[1].view(np.int32), rf.rules(0)[1].view(np.int32))
