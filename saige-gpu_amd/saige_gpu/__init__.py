"""saige_gpu: MI355X-native SAIGE step 1 hot path (libsaigehip) and its host
mirror of the reference's Rcpp exports and R drivers."""
from ._lib import SaigeError, NotConverged, load  # noqa: F401
from .api import Session  # noqa: F401
from . import driver, plink, synth  # noqa: F401
