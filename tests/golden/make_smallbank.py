"""Writes the word sequences of the reference's own test treebank
(src/test/resources/smallbank, Penn Treebank trees of Wall Street Journal text), one sentence per
line, tokens separated by spaces, in file order:
  smallbank_sentences.txt  dev, then test (68 sentences);
  smallbank_train15.txt    the first 100 train sentences of at most 15 tokens, the input of the
                           reference's CLParserTest (ParserTestHarness.getTrainTrees().take(100),
                           CLParserDataTest.scala:73-77, CLParserTest.scala:18-19,28-29).
These are real sentences for the text front end and the parse path; the trees themselves are not
used (the reference trains an epic grammar from them, which is not in the checkout).

Run from the repository root: python tests/golden/make_smallbank.py
"""
import os
import re

REF = "/root/reference/src/test/resources/smallbank"
HERE = os.path.dirname(os.path.abspath(__file__))


def trees(text):
    """Top-level bracketed trees of a treebank file."""
    out, depth, start = [], 0, 0
    for i, c in enumerate(text):
        if c == "(":
            if depth == 0:
                start = i
            depth += 1
        elif c == ")":
            depth -= 1
            if depth == 0:
                out.append(text[start:i + 1])
    return out


def words(tree):
    """The tree's yield: the word of every (tag word) leaf, empty elements (-NONE-) dropped."""
    return [w for tag, w in re.findall(r"\(([^\s()]+) ([^\s()]+)\)", tree) if tag != "-NONE-"]


def write(name, lines):
    path = os.path.join(HERE, name)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {len(lines)} sentences to {path}")


def main():
    lines = []
    for part in ("dev", "test"):
        with open(os.path.join(REF, part)) as f:
            lines += [" ".join(words(t)) for t in trees(f.read())]
    write("smallbank_sentences.txt", lines)
    with open(os.path.join(REF, "train")) as f:
        train = [words(t) for t in trees(f.read())]
    write("smallbank_train15.txt", [" ".join(w) for w in train if len(w) <= 15][:100])


if __name__ == "__main__":
    main()
