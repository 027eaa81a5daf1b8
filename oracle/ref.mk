# Builds oracle/_ref/ref_cdt from the reference's vendored CDT headers where
# they lie under /root/reference (never copied).  Checker only: used by
# tests/test_cpu_cdt.py in this container; absent elsewhere.
REF ?= /root/reference
_ref/ref_cdt: ref_cdt_main.cpp
	mkdir -p _ref
	g++ -O2 -std=c++17 -I$(REF)/include/CDT -o $@ ref_cdt_main.cpp
