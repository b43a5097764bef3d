"""Learned transmittance-PDT surrogate (device-trained MLP) and its data."""
from .model import DOMAIN, NB, PDTSurrogate, SurrogateEnsemble, encode, in_domain, ks_binned

__all__ = ["DOMAIN", "NB", "PDTSurrogate", "SurrogateEnsemble", "encode", "in_domain", "ks_binned"]
