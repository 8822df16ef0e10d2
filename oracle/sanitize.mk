# Builds the oracle with AddressSanitizer and UndefinedBehaviorSanitizer (host code only,
# g++) into a standalone driver, for tests/test_oracle_sanitizers.py.  TEST INFRASTRUCTURE.
CXX ?= g++
SANFLAGS := -O1 -g -fno-omit-frame-pointer -fsanitize=address,undefined -fno-sanitize-recover=undefined
CXXFLAGS := -std=c++17 -ffp-contract=off -fno-fast-math -fopenmp $(SANFLAGS)
BUILD := _build

all: $(BUILD)/sanitize_driver

$(BUILD)/sanitize_driver: sanitize_driver.cpp mahyco_oracle.cpp mahyco_oracle.h ../include/mahyco_types.h
	mkdir -p $(BUILD)
	$(CXX) $(CXXFLAGS) -o $@.tmp sanitize_driver.cpp mahyco_oracle.cpp && mv $@.tmp $@

.PHONY: all
