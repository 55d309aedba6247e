# Builds the real reference (bml 2.4.0 C sources, read in place under
# /root/reference) into oracle/_ref/ -- TEST INFRASTRUCTURE ONLY.
#
# The reference's own CMake build cannot run here (it always configures the
# Fortran interface and there is no gfortran), so this recipe compiles the
# C sources its CMake lists name for the C interface, dense, ellpack,
# ellblock, csr and internal-blas directories (distributed2d needs MPI and
# is left out), each *_typed.c once per precision as the reference's
# cmake/bmlAddTypedLibrary.cmake does.  No BLAS/LAPACK is available, so the
# build uses the reference's own NOBLAS configuration (bml_scale then reports
# "No BLAS library"); nothing is shimmed.  Outputs: oracle/_ref/libbmlref.so
# and oracle/_ref/ref_driver.
REF ?= /root/reference
SRC := $(REF)/src
OUT := _ref
CC := gcc
CFLAGS_REF := -O3 -fopenmp -std=gnu99 -fPIC -w -DBML_COMPLEX -DNOBLAS \
  -DMALLOC_ALIGNMENT=64 -DGLOBAL_DEBUG=BML_LOG_INFO -DBML_VERSION='"2.4.0"' \
  -D_POSIX_C_SOURCE=200112L -DBML_USE_POSIX_MEMALIGN \
  -I$(SRC)/C-interface -I$(SRC)
PRECS := SINGLE_REAL DOUBLE_REAL SINGLE_COMPLEX DOUBLE_COMPLEX

# source lists, read from the reference's CMakeLists.txt files
# (every *.c name inside each set(SOURCES-... ) list, the last one included,
# which shares its line with the closing parenthesis)
srclist = $(shell awk '/set\(SOURCES-/{f=1;next} f{for(i=1;i<=NF;i++) if ($$i ~ /[.]c/) {t=$$i; gsub(/[^A-Za-z0-9_.\/-]/,"",t); print t}} f&&/\)/{f=0}' $(1)/CMakeLists.txt)
DIRS := C-interface C-interface/dense C-interface/ellpack C-interface/ellblock \
        C-interface/csr internal-blas
ALLSRC := $(foreach d,$(DIRS),$(addprefix $(d)/,$(call srclist,$(SRC)/$(d))))
PLAIN := $(filter-out %_typed.c internal-blas/%,$(ALLSRC))
TYPED := $(filter %_typed.c internal-blas/%,$(ALLSRC))

objname = $(OUT)/obj/$(subst /,__,$(basename $(1)))
PLAIN_OBJ := $(foreach s,$(PLAIN),$(call objname,$(s)).o)
TYPED_OBJ := $(foreach s,$(TYPED),$(foreach p,$(PRECS),$(call objname,$(s))__$(p).o))

all: $(OUT)/libbmlref.so $(OUT)/ref_driver

define plain_rule
$(call objname,$(1)).o: $(SRC)/$(1) | $(OUT)/obj
	$(CC) $(CFLAGS_REF) -c $$< -o $$@
endef
define typed_rule
$(call objname,$(1))__$(2).o: $(SRC)/$(1) | $(OUT)/obj
	$(CC) $(CFLAGS_REF) -D$(2) -c $$< -o $$@
endef
$(foreach s,$(PLAIN),$(eval $(call plain_rule,$(s))))
$(foreach s,$(TYPED),$(foreach p,$(PRECS),$(eval $(call typed_rule,$(s),$(p)))))

$(OUT)/obj:
	mkdir -p $@

$(OUT)/libbmlref.so: $(PLAIN_OBJ) $(TYPED_OBJ)
	$(CC) -shared -fopenmp -o $@ $^ -lm

# bml_scale_inplace_csr_* is referenced by csr dispatch but compiled out
# under NOBLAS; it is never called, so the executable binds lazily.
$(OUT)/ref_driver: ref_driver.c $(OUT)/libbmlref.so
	$(CC) -O2 -std=gnu99 -fopenmp -Wall -Werror=implicit-function-declaration -Werror=int-conversion -I$(SRC)/C-interface -I$(SRC) -o $@ ref_driver.c \
	  -L$(OUT) -lbmlref -Wl,-rpath,'$$ORIGIN' -Wl,--allow-shlib-undefined -Wl,-z,lazy -lm

.PHONY: all
