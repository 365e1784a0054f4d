"""MI355X-native SQS scoring and search core (drop-in for sqsgenerator's optimize path).

Compute runs in libsqsgen_amd.so (HIP kernels for gfx950 behind a C ABI, declared in
include/sqsgen_amd.h); this package parses configurations and exposes results with the
reference's Python API shape.
"""
from .core import (ParseError, ParseErrorCode, SqsCallbackContext, SqsConfiguration, SqsResult,
                   SqsResultPack, Structure, distance_matrix, distances_histogram,
                   distances_naive, double, interact, num_permutations, optimize, parse_config,
                   prefactors, random, rank_permutation, shell_matrix, single, split, systematic,
                   unrank_permutation, Job, Server, batch_run, load_result_pack,
                   StoredResultPack)

__all__ = ["ParseError", "ParseErrorCode", "SqsCallbackContext", "SqsConfiguration", "SqsResult",
           "SqsResultPack", "Structure", "distance_matrix", "distances_histogram",
           "distances_naive", "double", "interact", "num_permutations", "optimize",
           "parse_config", "prefactors", "random", "rank_permutation", "shell_matrix", "single",
           "split", "systematic", "unrank_permutation", "Job", "Server", "batch_run",
           "load_result_pack", "StoredResultPack"]
