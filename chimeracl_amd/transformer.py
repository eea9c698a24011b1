"""Transformer: Fourier-Bessel transforms (chimeraCL/transformer.py:7-132).

The spectral axes, DHT matrices and solver tables are computed in FP64 on
the GPU (TransformerMethodsHIP._build_tables); host copies of kx, kr_m and
w_m are kept in Args as the reference does."""
from __future__ import annotations

from .methods.transformer_methods_hip import TransformerMethodsHIP


class Transformer(TransformerMethodsHIP):
    def init_transformer(self):
        """transformer.py:8-15."""
        self._init_transformer_data_on_dev()
        self.init_transformer_methods()
        self._build_tables()
