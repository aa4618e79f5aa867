"""The deterministic streams the reference's tests draw their inputs from, so the same inputs can be rebuilt.

bvh.rs:397-403 `stream(seed)`: s = s * 1664525 + 1013904223 (u32), a number (s >> 8) / 2^24 (f32).
units.rs:1245-1264 `Rng`: s = s * 6364136223846793005 + 1442695040888963407 (u64), a number (s >> 40) / 2^24.
"""
import numpy as np


class Stream:
    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFF

    def __call__(self) -> np.float32:
        self.s = (self.s * 1664525 + 1013904223) & 0xFFFFFFFF
        return np.float32(self.s >> 8) / np.float32(1 << 24)

    def p(self, s: float) -> np.ndarray:
        """V::new(rnd() - 0.5, rnd() - 0.5, rnd() - 0.5) * s, in f32."""
        x = self() - np.float32(0.5)
        y = self() - np.float32(0.5)
        z = self() - np.float32(0.5)
        return np.array([x, y, z], np.float32) * np.float32(s)


class Rng:
    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def f(self) -> np.float32:
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        return np.float32(self.s >> 40) / np.float32(1 << 24)

    def below(self, n: int) -> int:
        return min(int(self.f() * np.float32(n)), max(n - 1, 0))

    def v(self, s: float) -> np.ndarray:
        x = self.f() - np.float32(0.5)
        y = self.f() - np.float32(0.5)
        z = self.f() - np.float32(0.5)
        return np.array([x, y, z], np.float32) * np.float32(s)

    def tris(self, n: int, spread: float, size: float) -> np.ndarray:
        out = np.zeros((n, 3, 3), np.float32)
        for k in range(n):
            c = self.v(spread)
            out[k, 0] = c + self.v(size)
            out[k, 1] = c + self.v(size)
            out[k, 2] = c + self.v(size)
        return out
