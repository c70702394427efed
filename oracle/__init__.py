"""CPU oracle for the eSRRF path (test infrastructure; never imported by the product)."""
