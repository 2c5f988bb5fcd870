# Builds pieces of the reference itself, unchanged, from the files where they
# lie under $(REF) (read only): critic2's vendored LAPACK/BLAS (src/oldlibs,
# which matinv and eigsym call: tools_math@proc.f90:1203-1235, :1005-1039)
# and its FFT (src/cfftnd.f90, fftpack5).  Outputs go only into _ref/.
# Test infrastructure: tests/test_oracle_ref.py checks the oracle's
# restatements against this library.  Usage: make -f ref.mk [REF=...]
REF ?= /root/reference
FLANG ?= /opt/rocm/bin/amdflang
FFLAGS = -O3 -fPIC
OUT = _ref
LAPACK_SRC = $(wildcard $(REF)/src/oldlibs/*.f)
LAPACK_OBJ = $(patsubst $(REF)/src/oldlibs/%.f,$(OUT)/obj/%.o,$(LAPACK_SRC))
FFT_OBJ = $(OUT)/obj/cfftnd.o

all: $(OUT)/libc2ref.so

$(OUT)/obj:
	mkdir -p $@

$(OUT)/obj/%.o: $(REF)/src/oldlibs/%.f | $(OUT)/obj
	$(FLANG) $(FFLAGS) -c $< -o $@

$(OUT)/obj/cfftnd.o: $(REF)/src/cfftnd.f90 | $(OUT)/obj
	$(FLANG) $(FFLAGS) -c $< -o $@ -module-dir $(OUT)/obj

$(OUT)/libc2ref.so: $(LAPACK_OBJ) $(FFT_OBJ)
	$(FLANG) -shared -fPIC $^ -o $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
