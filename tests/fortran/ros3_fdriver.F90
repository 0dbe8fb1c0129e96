! Test driver for the Fortran binding (cmaq_amd/fortran/cmaq_ros3_c.F90): a
! Fortran caller as CCTM's ros3 CHEM would be, reading the RXNS_DATA tables
! and cell inputs from stream files that tests/test_fortran_binding.py writes.
!
!   ros3_fdriver sizes                           C_SIZEOF and offsets of the types
!   ros3_fdriver init  TABLES [NSPECIAL]         cmaq_ros3_init: return code, message
!   ros3_fdriver run   TABLES INPUTS OUT LOG     init + cmaq_ros3_run; outputs to OUT,
!                                                a failure dump to LOG
!   ros3_fdriver gear  TABLES INPUTS OUT         init + cmaq_gear_run; outputs to OUT
program ros3_fdriver
   use, intrinsic :: iso_c_binding
   use cmaq_ros3_c
   implicit none

   character(len=512) :: mode, ftab, fin, fout, flog, arg
   ! RXNS_DATA tables (TARGET copies, as the CHEM driver keeps them)
   integer(c_int32_t) :: ns, nr, mxprd, kunits, nfalloff, nmphot, nphotab, mhetero, nhetero, nspecial
   integer(c_int32_t) :: nw(6)
   integer(c_int32_t), allocatable, target :: irr(:, :), nreact(:), nprdct(:), ktype(:), irxbits(:), &
      iorder(:), irrfall(:), iph(:, :), ihetero(:, :), w1(:), w2(:), w3(:), w4(:), w5(:), w6(:)
   real(c_double), allocatable, target :: sc(:, :), rtdat(:, :), rfdat(:, :), rtol(:), atol(:)
   real(c_double) :: atm(5)
   character(len=16), allocatable :: spc(:), photab(:), rxlabel(:)
   type(cmaq_ros3_mech_t) :: mech
   type(cmaq_ros3_opts_t) :: opts
   type(c_ptr) :: ctx
   integer(c_int) :: rc

   call get_command_argument(1, mode)
   select case (trim(mode))
   case ('sizes')
      call print_sizes()
   case ('init')
      call get_command_argument(2, ftab)
      call read_tables(ftab)
      nspecial = 0
      if (command_argument_count() >= 3) then
         call get_command_argument(3, arg)
         read (arg, *) nspecial
      end if
      call fill_mech()
      rc = cmaq_ros3_init(ctx, mech, rtol, atol, opts)
      write (*, '(a, i0)') 'rc ', rc
      write (*, '(a, a)') 'message ', cmaq_ros3_error_message(ctx)
      if (rc == CMAQ_ROS3_OK) call cmaq_ros3_free(ctx)
   case ('run')
      call get_command_argument(2, ftab)
      call get_command_argument(3, fin)
      call get_command_argument(4, fout)
      call get_command_argument(5, flog)
      call read_tables(ftab)
      nspecial = 0
      call fill_mech()
      rc = cmaq_ros3_init(ctx, mech, rtol, atol, opts)
      if (rc /= CMAQ_ROS3_OK) then
         write (*, '(a, i0, 1x, a)') 'init failed ', rc, cmaq_ros3_error_message(ctx)
         stop 2
      end if
      call run_cells(fin, fout, flog)
   case ('gear')
      call get_command_argument(2, ftab)
      call get_command_argument(3, fin)
      call get_command_argument(4, fout)
      call read_tables(ftab)
      nspecial = 0
      call fill_mech()
      rc = cmaq_ros3_init(ctx, mech, rtol, atol, opts)
      if (rc /= CMAQ_ROS3_OK) then
         write (*, '(a, i0, 1x, a)') 'init failed ', rc, cmaq_ros3_error_message(ctx)
         stop 2
      end if
      call gear_cells(fin, fout)
      call cmaq_ros3_free(ctx)
   case default
      write (*, '(a)') 'usage: ros3_fdriver sizes | init TABLES [NSPECIAL] | run TABLES INPUTS OUT LOG | gear TABLES INPUTS OUT'
      stop 1
   end select

contains

   integer(c_intptr_t) function addr(p)
      type(c_ptr), intent(in) :: p
      addr = transfer(p, 0_c_intptr_t)
   end function addr

   subroutine print_sizes()
      type(cmaq_ros3_mech_t), target :: m
      type(cmaq_ros3_opts_t), target :: o
      type(cmaq_ros3_fail_t), target :: f
      type(cmaq_ros3_fail_detail_t), target :: d
      type(cmaq_ros3_grid_t), target :: g
      type(cmaq_gear_tol_t), target :: t
      integer(c_intptr_t) :: b
      write (*, '(a, i0)') 'sizeof mech ', c_sizeof(m)
      write (*, '(a, i0)') 'sizeof opts ', c_sizeof(o)
      write (*, '(a, i0)') 'sizeof fail ', c_sizeof(f)
      write (*, '(a, i0)') 'sizeof fail_detail ', c_sizeof(d)
      write (*, '(a, i0)') 'sizeof grid ', c_sizeof(g)
      write (*, '(a, i0)') 'sizeof gear_tol ', c_sizeof(t)
      b = addr(c_loc(m))
      write (*, '(a, i0)') 'offset mech.nrxns ', addr(c_loc(m%nrxns)) - b
      write (*, '(a, i0)') 'offset mech.irr ', addr(c_loc(m%irr)) - b
      write (*, '(a, i0)') 'offset mech.rtdat ', addr(c_loc(m%rtdat)) - b
      write (*, '(a, i0)') 'offset mech.nfalloff ', addr(c_loc(m%nfalloff)) - b
      write (*, '(a, i0)') 'offset mech.irrfall ', addr(c_loc(m%irrfall)) - b
      write (*, '(a, i0)') 'offset mech.nmphot ', addr(c_loc(m%nmphot)) - b
      write (*, '(a, i0)') 'offset mech.iph ', addr(c_loc(m%iph)) - b
      write (*, '(a, i0)') 'offset mech.mhetero ', addr(c_loc(m%mhetero)) - b
      write (*, '(a, i0)') 'offset mech.ihetero ', addr(c_loc(m%ihetero)) - b
      write (*, '(a, i0)') 'offset mech.nwm ', addr(c_loc(m%nwm)) - b
      write (*, '(a, i0)') 'offset mech.nrxwm ', addr(c_loc(m%nrxwm)) - b
      write (*, '(a, i0)') 'offset mech.nrxwh2 ', addr(c_loc(m%nrxwh2)) - b
      write (*, '(a, i0)') 'offset mech.atm_air ', addr(c_loc(m%atm_air)) - b
      write (*, '(a, i0)') 'offset mech.atm_h2 ', addr(c_loc(m%atm_h2)) - b
      write (*, '(a, i0)') 'offset mech.nspecial_rxn ', addr(c_loc(m%nspecial_rxn)) - b
      b = addr(c_loc(o))
      write (*, '(a, i0)') 'offset opts.refill_min ', addr(c_loc(o%refill_min)) - b
      b = addr(c_loc(f))
      write (*, '(a, i0)') 'offset fail.cell ', addr(c_loc(f%cell)) - b
      write (*, '(a, i0)') 'offset fail.tnow ', addr(c_loc(f%tnow)) - b
      b = addr(c_loc(d))
      write (*, '(a, i0)') 'offset fail_detail.cell ', addr(c_loc(d%cell)) - b
      write (*, '(a, i0)') 'offset fail_detail.err ', addr(c_loc(d%err)) - b
      write (*, '(a, i0)') 'offset fail_detail.dens ', addr(c_loc(d%dens)) - b
      b = addr(c_loc(g))
      write (*, '(a, i0)') 'offset grid.cgrid ', addr(c_loc(g%cgrid)) - b
      write (*, '(a, i0)') 'offset grid.khetero ', addr(c_loc(g%khetero)) - b
      b = addr(c_loc(t))
      write (*, '(a, i0)') 'offset gear_tol.conc_floor ', addr(c_loc(t%conc_floor)) - b
   end subroutine print_sizes

   ! Stream file written by the test: sizes, then each table in Fortran order.
   subroutine read_tables(path)
      character(len=*), intent(in) :: path
      integer :: u
      open (newunit=u, file=trim(path), access='stream', form='unformatted', status='old')
      read (u) ns, nr, mxprd, kunits, nfalloff, nmphot, nphotab, mhetero, nhetero, nw
      allocate (irr(nr, mxprd + 3), sc(nr, mxprd), nreact(nr), nprdct(nr), ktype(nr), irxbits(nr), &
         iorder(nr), rtdat(3, nr), irrfall(max(nfalloff, 1)), rfdat(5, max(nfalloff, 1)), &
         iph(max(nmphot, 1), 3), ihetero(max(mhetero, 1), 2), w1(max(nw(1), 1)), w2(max(nw(2), 1)), &
         w3(max(nw(3), 1)), w4(max(nw(4), 1)), w5(max(nw(5), 1)), w6(max(nw(6), 1)), rtol(ns), atol(ns), &
         spc(ns), photab(max(nphotab, 1)), rxlabel(nr))
      read (u) irr, sc, nreact, nprdct, ktype, irxbits, iorder, rtdat, irrfall, rfdat, iph, ihetero
      read (u) w1, w2, w3, w4, w5, w6, atm, rtol, atol
      read (u) spc, photab, rxlabel
      close (u)
   end subroutine read_tables

   subroutine fill_mech()
      mech%numb_mech_spc = ns
      mech%nrxns = nr
      mech%mxprd = mxprd
      mech%kunits = kunits
      mech%irr = c_loc(irr)
      mech%sc = c_loc(sc)
      mech%nreact = c_loc(nreact)
      mech%nprdct = c_loc(nprdct)
      mech%ktype = c_loc(ktype)
      mech%irxbits = c_loc(irxbits)
      mech%iorder = c_loc(iorder)
      mech%rtdat = c_loc(rtdat)
      mech%nfalloff = nfalloff
      mech%irrfall = c_loc(irrfall)
      mech%rfdat = c_loc(rfdat)
      mech%nmphot = nmphot
      mech%nphotab = nphotab
      mech%iph = c_loc(iph)
      mech%mhetero = mhetero
      mech%nhetero = nhetero
      mech%ihetero = c_loc(ihetero)
      mech%nwm = nw(1)
      mech%nww = nw(2)
      mech%nwo2 = nw(3)
      mech%nwn2 = nw(4)
      mech%nwch4 = nw(5)
      mech%nwh2 = nw(6)
      mech%nrxwm = c_loc(w1)
      mech%nrxww = c_loc(w2)
      mech%nrxwo2 = c_loc(w3)
      mech%nrxwn2 = c_loc(w4)
      mech%nrxwch4 = c_loc(w5)
      mech%nrxwh2 = c_loc(w6)
      mech%atm_air = atm(1)
      mech%atm_o2 = atm(2)
      mech%atm_n2 = atm(3)
      mech%atm_ch4 = atm(4)
      mech%atm_h2 = atm(5)
      mech%nspecial_rxn = nspecial
   end subroutine fill_mech

   ! One chemistry step over the file's cells, as CHEM's block loop would:
   ! cell-fastest block arrays in, Y and the step counts out; a convergence
   ! failure is written to LOG the way RBSOLVER writes it before M3EXIT.
   subroutine run_cells(pin, pout, plog)
      character(len=*), intent(in) :: pin, pout, plog
      integer(c_int64_t) :: n
      real(c_double) :: chemstep
      real(c_double), allocatable :: temp(:), pres(:), h2o(:), rj(:, :), y(:, :)
      integer(c_int32_t), allocatable :: nacc(:), nrej(:)
      type(cmaq_ros3_fail_t) :: fail
      integer :: u, code
      open (newunit=u, file=trim(pin), access='stream', form='unformatted', status='old')
      read (u) n, chemstep
      allocate (temp(n), pres(n), h2o(n), rj(n, max(nphotab, 1)), y(n, ns), nacc(n), nrej(n))
      read (u) temp, pres, h2o, rj, y
      close (u)
      rc = cmaq_ros3_run(ctx, n, chemstep, temp, pres, h2o, rj, c_null_ptr, c_null_ptr, y, nacc, nrej, fail)
      write (*, '(a, i0, a, i0)') 'rc ', rc, ' cell ', fail%cell
      if (rc == CMAQ_ROS3_E_CONVERGENCE .or. rc == CMAQ_ROS3_E_STEPS) then
         open (newunit=u, file=trim(plog), status='replace', action='write')
         code = cmaq_ros3_write_failure(u, ctx, 2018001, 120000, spc, photab(1:nphotab), rxlabel, &
            rtdat, iph)
         close (u)
         write (*, '(a, i0)') 'dump ', code
      end if
      open (newunit=u, file=trim(pout), access='stream', form='unformatted', status='replace')
      write (u) rc, y, nacc, nrej
      close (u)
   end subroutine run_cells

   ! The same cells through SMVGEAR with the default Gear tolerances.
   subroutine gear_cells(pin, pout)
      character(len=*), intent(in) :: pin, pout
      integer(c_int64_t) :: n
      real(c_double) :: chemstep
      real(c_double), allocatable :: temp(:), pres(:), h2o(:), rj(:, :), y(:, :)
      integer(c_int32_t), allocatable :: nsteps(:), nfail(:)
      type(cmaq_ros3_fail_t) :: fail
      type(cmaq_gear_tol_t) :: tol
      integer :: u
      open (newunit=u, file=trim(pin), access='stream', form='unformatted', status='old')
      read (u) n, chemstep
      allocate (temp(n), pres(n), h2o(n), rj(n, max(nphotab, 1)), y(n, ns), nsteps(n), nfail(n))
      read (u) temp, pres, h2o, rj, y
      close (u)
      rc = cmaq_gear_set_tolerances(ctx, tol)
      if (rc == CMAQ_ROS3_OK) &
         rc = cmaq_gear_run(ctx, n, chemstep, temp, pres, h2o, rj, c_null_ptr, c_null_ptr, y, nsteps, nfail, fail)
      write (*, '(a, i0, a, i0)') 'rc ', rc, ' cell ', fail%cell
      open (newunit=u, file=trim(pout), access='stream', form='unformatted', status='replace')
      write (u) rc, y, nsteps, nfail
      close (u)
   end subroutine gear_cells

end program ros3_fdriver
