//go:build linux && amd64

package rocm_test

// Properties of the reference's own tests (api_test.go, compressor_test.go,
// format_conformance_test.go:727-754) checked through the GPU binding:
// the golden blocks byte for byte, round trips of every flag combination,
// Get equal to Unpack at every position, BlockLength equal to the bytes
// consumed, the reference's error values on empty and short buffers, and a
// Compressor reused across rounds. The golden files are the reference's
// testdata, kept in this repository under tests/golden.

import (
	"errors"
	"os"
	"path/filepath"
	"slices"
	"testing"

	"github.com/Akron/utlpfor"
	"github.com/Akron/utlpfor/rocm"
)

func golden(t *testing.T, name string) []byte {
	t.Helper()
	b, err := os.ReadFile(filepath.Join("..", "..", "tests", "golden", name))
	if err != nil {
		t.Fatal(err)
	}
	return b
}

func TestGoldenBlocks(t *testing.T) {
	plain := make([]uint32, 128)
	delta := make([]uint32, 128)
	exc := make([]uint32, 128)
	for i := range plain {
		plain[i] = uint32(128 + i)
		delta[i] = uint32(100 * i)
		exc[i] = uint32(i)
	}
	for p := 5; p < 100; p += 10 {
		exc[p] = 0x10000000 + uint32(p)
	}
	cases := []struct {
		file   string
		flag   rocm.Flag
		values []uint32
	}{
		{"plain_128_bw8.bin", 0, plain},
		{"plain_128_bw0.bin", 0, make([]uint32, 128)},
		{"delta_128.bin", utlpfor.Delta, delta},
		{"exceptions_128.bin", 0, exc},
	}
	for _, c := range cases {
		want := golden(t, c.file)
		got, err := rocm.PackUint32(c.flag, slices.Clone(c.values), nil, nil)
		if err != nil || !slices.Equal(got, want) {
			t.Fatalf("%s: %x, %v", c.file, got, err)
		}
		vals, consumed, err := rocm.UnpackUint32(want, nil, nil)
		if err != nil || consumed != len(want) || !slices.Equal(vals, c.values) {
			t.Fatalf("%s: unpack %v %d %v", c.file, vals, consumed, err)
		}
	}
}

func TestRoundTripEveryFlag(t *testing.T) {
	for flag := rocm.Flag(0); flag < 16; flag++ {
		for _, n := range []int{1, 5, 64, 128} {
			values := make([]uint32, n)
			for i := range values {
				values[i] = uint32(1000 + 7*i)
				if i%9 == 4 {
					values[i] = 0xFFFFFF00
				}
			}
			orig := slices.Clone(values)
			packed, err := rocm.PackUint32(flag, values, nil, nil)
			if err != nil {
				t.Fatal(flag, n, err)
			}
			if bl, err := rocm.BlockLength(packed); err != nil || bl != len(packed) {
				t.Fatal(flag, n, bl, err)
			}
			got, consumed, err := rocm.UnpackUint32(packed, nil, nil)
			if err != nil || consumed != len(packed) || !slices.Equal(got, orig) {
				t.Fatal(flag, n, err)
			}
			for i, want := range got {
				if v, err := rocm.GetUint32(i, packed, nil); err != nil || v != want {
					t.Fatal(flag, n, i, v, err)
				}
			}
		}
	}
}

func TestErrorsAreTheReferenceValues(t *testing.T) {
	if _, err := rocm.BlockLength([]byte{1, 2}); !errors.Is(err, utlpfor.ErrInvalidBuffer) {
		t.Fatal(err)
	}
	if _, _, err := rocm.UnpackUint32(nil, nil, nil); !errors.Is(err, utlpfor.ErrInvalidBuffer) {
		t.Fatal(err)
	}
	packed, _ := rocm.PackUint32(0, []uint32{1, 2, 3}, nil, nil)
	if _, err := rocm.GetUint32(3, packed, nil); !errors.Is(err, utlpfor.ErrPositionOutOfRange) {
		t.Fatal(err)
	}
}

func TestCompressorReuse(t *testing.T) {
	c := rocm.NewUint32()
	for round := 0; round < 5; round++ {
		values := make([]uint32, 32)
		for i := range values {
			values[i] = uint32(round*1000 + i)
		}
		orig := slices.Clone(values)
		packed, err := c.Compress(utlpfor.Delta, nil, values)
		if err != nil {
			t.Fatal(err)
		}
		got, _, err := c.Decompress(nil, packed)
		if err != nil || !slices.Equal(got, orig) {
			t.Fatal(round, err)
		}
	}
	c64 := rocm.NewUint64()
	v64 := []uint64{1_000_000_000_000, 1_000_000_000_001, 1_000_000_000_002}
	packed, err := c64.Compress(0, nil, v64)
	if err != nil || len(packed) != 19 { // example_test.go:206-216
		t.Fatal(len(packed), err)
	}
}
