"""The reference tests' LCG (seed = seed * 1664525 + 1013904223 mod 2^32),
e.g. crates/rustmpeg-video/src/vvc/dct.rs:1341-1345."""


class Lcg:
    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFF

    def next_u32(self):
        self.s = (self.s * 1664525 + 1013904223) & 0xFFFFFFFF
        return self.s
