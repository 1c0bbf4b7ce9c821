"""MI355X-native Relay-BP-S decoder (drop-in for the reference's RelayBPDecoder).

The decoding arithmetic runs in hand-written HIP kernels for gfx950
(``csrc/relaybp.hip``) reached through the C ABI of
``include/pelagos_relaybp.h``; this package is the host side.
"""

from .decoder import RelayBPDecoder

__all__ = ["RelayBPDecoder"]
