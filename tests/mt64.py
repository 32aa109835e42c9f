"""MT19937-64 written from its published algorithm, for tests that pin sampling draws."""
import numpy as np


class Mt64:
    """MT19937-64 (Matsumoto & Nishimura, ACM TOMACS 2000), written from the published algorithm:
    the generator behind std::mt19937_64, to pin the oracle's draw sequence independently."""
    NN, MM, MATRIX_A, UM, LM = 312, 156, 0xB5026F5AA96619E9, 0xFFFFFFFF80000000, 0x7FFFFFFF
    MASK = (1 << 64) - 1

    def __init__(self, seed=5489):
        self.mt = [0] * self.NN
        self.mt[0] = seed & self.MASK
        for i in range(1, self.NN):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & self.MASK
        self.i = self.NN

    def next(self):
        if self.i >= self.NN:
            mt = self.mt
            for k in range(self.NN):
                x = (mt[k] & self.UM) | (mt[(k + 1) % self.NN] & self.LM)
                xa = x >> 1
                if x & 1:
                    xa ^= self.MATRIX_A
                mt[k] = mt[(k + self.MM) % self.NN] ^ xa
            self.i = 0
        x = self.mt[self.i]
        self.i += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & self.MASK

    def uniform(self):
        """std::uniform_real_distribution<double>(0, 1): generate_canonical with one 64-bit draw."""
        r = float(self.next()) / 18446744073709551616.0
        return r if r < 1.0 else np.nextafter(1.0, 0.0)
