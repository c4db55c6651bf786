"""Continuous batching of CW solves (mixed grid sizes, HBM admission, batched launches)."""

from caustica_amd.serving.batcher import (
    ContinuousBatcher,
    SolveRequest,
    SolveResponse,
    replay_fifo,
    result_digest,
    serve_serial,
    summarize,
)

__all__ = ["ContinuousBatcher", "SolveRequest", "SolveResponse", "replay_fifo", "result_digest",
           "serve_serial", "summarize"]
