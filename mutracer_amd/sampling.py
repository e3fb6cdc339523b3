"""Host-side sampling parameters (src/renderer/sampling.ts:1-34)."""
import math


def sampling_pattern_code(pattern):
    codes = {"sobol": 0, "zsobol": 1, "random": 2}
    if pattern not in codes:
        raise ValueError(f"Unknown sampling pattern: {pattern}")
    return codes[pattern]


def zsobol_log2spp(max_samples):
    """256^2 pixel tiles leave 16 bits for samples in a 32-bit Morton index."""
    if not math.isfinite(max_samples):
        return 16
    return min(16, max(0, math.ceil(math.log2(max(1, max_samples)))))


def trace_sampling(pattern, max_samples, sequence_seed):
    """sceneAssembly.ts:32-35: ZSobol reserves sample 0, so its budget is max - 1."""
    code = sampling_pattern_code(pattern)
    log = zsobol_log2spp(max_samples - 1 if code == 1 else max_samples)
    return code, log, sequence_seed & 0xFFFFFFFF
