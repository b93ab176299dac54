"""Mixture constant pipeline: species data, fits, eos, transport and chemistry tables."""

from peregrine_amd.mixture.constants import Ru, Tref, kB, avogadro, pref
from peregrine_amd.mixture.mixture import Mixture, get_mixture
from peregrine_amd.mixture.species import Species

__all__ = ["Mixture", "Ru", "Species", "Tref", "avogadro", "get_mixture", "kB", "pref"]
