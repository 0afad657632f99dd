"""MI355X-native pseudo-particle self-energy engine (drop-in for triqs_xca's DiagramEvaluator path)."""
from .diag import all_connected_pairings, all_pairings  # noqa: F401
