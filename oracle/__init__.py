"""CPU oracle of the reference CPA integration path (test infrastructure only)."""
