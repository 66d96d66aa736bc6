module github.com/Akron/utlpfor/rocm

go 1.26

require github.com/Akron/utlpfor v0.0.0

// A maintainer's checkout of the reference module next to this repository.
replace github.com/Akron/utlpfor => ../../../utlpfor
