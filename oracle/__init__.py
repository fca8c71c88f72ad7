"""Test-only checkers: FP64 C restatement of the reference path and the
reference's own C built from its sources.  Never imported by the product."""
