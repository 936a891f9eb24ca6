"""One HIP stream per parser, and no cross-stream waits anywhere in the engine (CPU check of the
sources).

Parsers of one process share the GPU's hardware queues (GPU_MAX_HW_QUEUES, four by default): with
more streams than queues, several streams feed one in-order queue. A wait on another stream's
event is a barrier packet in that queue, and it holds back every packet queued behind it, including
other streams' packets. Two parsers that each wait across their own two streams can then form a
cycle through two shared queues: parser A's wait sits in queue 1 ahead of parser B's plan kernels,
while B's wait sits in queue 2 ahead of A's. Keeping each parser's work on its own single stream
(mask plans included) leaves no cross-stream wait, hence no such cycle, however many parsers run.
"""
import os
import re

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "puck_amd", "csrc")


def _sources():
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hip", ".cpp", ".h")):
            with open(os.path.join(CSRC, name)) as f:
                yield name, re.sub(r"//[^\n]*", "", f.read())


def test_no_cross_stream_waits():
    for name, text in _sources():
        assert "hipStreamWaitEvent" not in text, name


def test_one_stream_per_parser():
    creates = [(name, len(re.findall(r"hipStreamCreate\w*\s*\(", text))) for name, text in _sources()]
    assert [c for c in creates if c[1]] == [("engine.hip", 1)], creates
    engine = dict(_sources())["engine.hip"]
    ctor = engine[engine.index("Parser::Parser("):engine.index("Parser::~Parser(")]
    assert re.search(r"hipStreamCreate\w*\s*\(&stream_", ctor)
