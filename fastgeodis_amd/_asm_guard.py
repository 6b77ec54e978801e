"""Build-time guard: lists scalar-data-cache write instructions (scalar
stores/atomics and the scalar cache's write-back/discard) found in device
assembly.  The library writes only through vector memory instructions; the
build fails if the compiler ever emits one of these.  Host-side only."""
import re

_PATTERN = re.compile(
    r"^\s*(s_store_\w+|s_buffer_store\w*|s_scratch_store\w*|s_atomic_\w+|s_buffer_atomic_\w+|"
    r"s_dcache_wb\w*|s_dcache_discard\w*)", re.M)


def scalar_cache_writes(asm_text: str):
    return sorted(set(_PATTERN.findall(asm_text)))
