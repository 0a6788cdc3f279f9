"""ORACLE — test infrastructure only (CPU restatement of the reference path and the compiled-reference harness)."""
